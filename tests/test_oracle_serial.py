"""The serial oracle (oracle/serial.c) against the reference's own checks.

Each test restates one of glazier-core's or glazier-tests' CPU tests: the
incremental energy change against a full recomputation, counters against
recounts, reproducibility, the behavioural thresholds, the moments and field
known answers, and the connectivity test against brute force.
"""

import itertools

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model, Species, default_cell_types, tiled_labels


def base_model(w=32, h=32, d=1, seed=42, **cell):
    """cpu_reference.rs:7-29: medium-cell 16, cell-cell 8, V_t 64, lambda_V 2."""
    c = CellType(target_volume=64.0, lambda_volume=2.0, **cell)
    return Model(width=w, height=h, depth=d, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), c], temperature=10.0, neighbour_order=2, seed=seed)


def tiled(model, side=8, nx=None, ny=None, nz=1, types=None):
    nx = model.width // side if nx is None else nx
    ny = model.height // side if ny is None else ny
    labels = tiled_labels(model, side, nx, ny, nz)
    n = int(labels.max())
    return oracle.SerialSimulation(model, labels, default_cell_types(n + 1, types))


def check_delta_against_energy(sim, rng, steps, probes):
    for _ in range(steps):
        sim.step(1)
        for _ in range(probes):
            target = rng.below(sim.sites)
            source = rng.below(sim.sites)
            labels = sim.labels()
            new, old = int(labels[source]), int(labels[target])
            vol = sim.volumes()
            if old == new or (old != 0 and vol[old] <= 1):
                continue
            predicted = sim.delta_energy(target, new)
            before = sim.energy()
            after = sim.clone()
            after.apply_copy(target, new)
            measured = after.energy() - before
            after.close()
            assert abs(predicted - measured) < 1e-9, (predicted, measured, target)


def test_incremental_delta_matches_full_recomputation(oracle_lib):
    # cpu_reference.rs:32-74
    check_delta_against_energy(tiled(base_model()), oracle.Xoshiro(5), 10, 20)


def test_incremental_delta_with_surface_term(oracle_lib):
    # cpu_reference.rs:85-119 (target 28, lambda_S 1)
    m = base_model(target_surface=28.0, lambda_surface=1.0)
    check_delta_against_energy(tiled(m), oracle.Xoshiro(9), 8, 20)


def test_incremental_delta_with_length_term(oracle_lib):
    # length.rs:30-71 (target 20, lambda_L 1, 48x48, seed 3)
    m = base_model(48, 48, seed=3, target_length=20.0, lambda_length=1.0)
    check_delta_against_energy(tiled(m), oracle.Xoshiro(2), 6, 15)


def test_bookkeeping_survives_a_run(oracle_lib):
    # cpu_reference.rs:121-128, 139-146
    sim = tiled(base_model(48, 48, target_surface=28.0, lambda_surface=1.0))
    sim.step(100)
    assert (sim.surfaces() == sim.recounted_surfaces()).all()
    assert (sim.volumes() == sim.recounted_volumes()).all()


def test_no_cell_is_lost(oracle_lib):
    # cpu_reference.rs:148-161
    sim = tiled(base_model(48, 48))
    n = sim.n_labels() - 1
    sim.step(200)
    assert sim.n_labels() - 1 == n
    assert (sim.recounted_volumes()[1:] >= 1).all()


def test_a_seed_reproduces_a_run(oracle_lib):
    # cpu_reference.rs:163-172
    a, b = tiled(base_model()), tiled(base_model())
    a.step(25)
    b.step(25)
    assert (a.labels() == b.labels()).all()


def test_surface_constraint_pulls_the_outline_in(oracle_lib):
    # cpu_reference.rs:130-137
    loose = tiled(base_model(48, 48))
    tight = tiled(base_model(48, 48, target_surface=28.0, lambda_surface=1.0))
    loose.step(200)
    tight.step(200)
    assert tight.recounted_surfaces()[1:].mean() < loose.recounted_surfaces()[1:].mean()


def test_volume_constraint_pulls_cells_to_target(oracle_lib):
    # cpu_reference.rs:174-190: cells of 36 sites grow toward 64
    m = base_model(64, 64)
    sim = tiled(m, side=6, nx=6, ny=6)
    start = sim.volumes()[1:].mean()
    sim.step(300)
    end = sim.volumes()[1:].mean()
    assert abs(start - 36.0) < 1e-9
    assert start < end < 72.0


def test_moments_known_answers(oracle_lib):
    # moments.rs:204-285
    def filled(w, h, d):
        pts = [(x, y, z) for z in range(d) for y in range(h) for x in range(w)]
        return np.array(pts, dtype=np.float64).reshape(-1)

    def length(w, h, d):
        a = filled(w, h, d)
        return oracle_lib.oracle_moments_length(a.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)), a.size // 3)

    # 1 x 16 strip: 4 sqrt((16^2 - 1)/12) = 18.439...
    assert abs(length(1, 16, 1) - 4 * np.sqrt((16 * 16 - 1) / 12)) < 1e-9
    assert abs(length(1, 16, 1) - 18.47) < 0.05
    assert length(16, 4, 1) > length(8, 8, 1)
    assert abs(length(16, 4, 1) - length(4, 16, 1)) < 1e-6
    assert abs(length(1, 1, 16) - length(1, 16, 1)) < 1e-6
    # a square: the closed form of a uniform k x k grid, 4 sqrt((k^2-1)/12)
    assert abs(length(8, 8, 1) - 4 * np.sqrt(63 / 12)) < 1e-9


def test_principal_axis_matches_numpy(oracle_lib):
    import ctypes

    rng = np.random.default_rng(3)
    for _ in range(50):
        a = rng.normal(size=(3, 3))
        c = a @ a.T
        ax = (ctypes.c_double * 3)()
        lam = ctypes.c_double()
        cc = c.reshape(-1).copy()
        oracle_lib.oracle_principal_axis(cc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ax,
                                         ctypes.byref(lam))
        w, v = np.linalg.eigh(c)
        gap = (w[2] - w[1]) / w[2]
        if gap < 0.2:
            continue  # twenty rounds only converge with a spectral gap
        assert abs(lam.value - w[2]) / w[2] < 1e-6


def field_model(w, h, d, diffusion, decay, initial):
    m = Model(width=w, height=h, depth=d, contact=[0.0, 16.0, 16.0, 8.0],
              types=[CellType(), CellType(target_volume=64.0, lambda_volume=2.0)],
              species=[Species("a", diffusion, decay, initial)],
              secretion=[[0.0], [0.0]], uptake=[[0.0], [0.0]], chemotaxis=[[0.0], [0.0]])
    return m


def medium_sim(m, values=None):
    labels = np.zeros(m.sites, dtype=np.uint32)
    sim = oracle.SerialSimulation(m, labels, np.ones(1, dtype=np.uint8))
    if values is not None:
        sim.set_field(0, values)
    return sim


def test_uniform_field_stays_uniform(oracle_lib):
    # field.rs:177-192
    sim = medium_sim(field_model(16, 16, 1, 0.2, 0.0, 3.0))
    sim.fields_step()
    assert np.abs(sim.field(0) - 3.0).max() < 1e-12


def test_diffusion_conserves_the_total(oracle_lib):
    # field.rs:194-216 (plane) and 284-307 (volume)
    for (w, h, d, dcoef, centre, steps) in ((16, 16, 1, 0.2, 8 * 16 + 8, 50),
                                            (12, 12, 12, 0.15, (6 * 12 + 6) * 12 + 6, 40)):
        v = np.zeros(w * h * d)
        v[centre] = 100.0
        sim = medium_sim(field_model(w, h, d, dcoef, 0.0, 0.0), v)
        for _ in range(steps):
            sim.fields_step()
        assert abs(sim.field(0).sum() - 100.0) < 1e-9


def test_decay_removes_a_stated_fraction(oracle_lib):
    # field.rs:236-249
    sim = medium_sim(field_model(8, 8, 1, 0.0, 0.1, 1.0))
    sim.fields_step()
    assert abs(sim.field(0)[0] - 0.9) < 1e-12


def test_substep_table(oracle_lib):
    # field.rs:251-268
    for (dcoef, dims, want) in ((0.0, 2, 1), (0.25, 2, 1), (0.26, 2, 2), (2.0, 2, 8),
                                (1.0, 2, 4), (1.0, 3, 6)):
        m = field_model(4, 4, 4 if dims == 3 else 1, dcoef, 0.0, 0.0)
        assert oracle_lib.serial_substeps(0, __import__("ctypes").byref(m.to_c())) == want


def test_secretion_then_uptake(oracle_lib):
    # field.rs:326-345: owned site 1 + 2 = 3, less half = 1.5
    m = field_model(8, 8, 1, 0.0, 0.0, 1.0)
    m.secretion = [[0.0], [2.0]]
    m.uptake = [[0.0], [0.5]]
    labels = np.zeros(64, dtype=np.uint32)
    labels[0] = 1
    sim = oracle.SerialSimulation(m, labels, np.array([0, 1], dtype=np.uint8))
    sim.fields_step()
    f = sim.field(0)
    assert abs(f[0] - 1.5) < 1e-12 and abs(f[1] - 1.0) < 1e-12


def ring_positions(ring):
    return [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dx, dy, dz) != (0, 0, 0) and (ring == 26 or dz == 0)]


def bfs_one_piece(members):
    if len(members) <= 1:
        return True
    seen = {members[0]}
    stack = [members[0]]
    while stack:
        a = stack.pop()
        for b in members:
            if b not in seen and all(abs(p - q) <= 1 for p, q in zip(a, b)):
                seen.add(b)
                stack.append(b)
    return len(seen) == len(members)


def test_planar_connectivity_every_mask(oracle_lib):
    # connectivity.rs:43-84 on every one of the 2^8 planar rings, against an
    # independent breadth-first search, through the serial engine itself
    pos = ring_positions(8)
    m = Model(width=9, height=9, depth=1, contact=[0.0, 1.0, 1.0, 1.0],
              types=[CellType(), CellType(connected=True)], seed=1)
    for mask in range(256):
        members = [p for k, p in enumerate(pos) if mask >> k & 1]
        labels = np.zeros((9, 9), dtype=np.uint32)
        for (dx, dy, _) in members:
            labels[4 + dy, 4 + dx] = 1
        sim = oracle.SerialSimulation(m, labels.reshape(-1), np.array([1, 1], dtype=np.uint8))
        assert sim.locally_connected(4 * 9 + 4, 1) == bfs_one_piece(members), mask
        assert oracle.one_piece(mask, 8) == bfs_one_piece(members), mask


def test_volume_connectivity_sampled_masks(oracle_lib):
    pos = ring_positions(26)
    rng = np.random.default_rng(0)
    masks = list(rng.integers(0, 1 << 26, size=3000)) + [0, 1, (1 << 26) - 1]
    masks += [(1 << 4) | (1 << 21)]  # the bridge through z, connectivity.rs:134-151
    for mask in masks:
        members = [p for k, p in enumerate(pos) if int(mask) >> k & 1]
        assert oracle.one_piece(int(mask), 26) == bfs_one_piece(members), mask


def test_connectivity_reference_cases(oracle_lib):
    # connectivity.rs:99-152
    pos = ring_positions(8)
    idx = {p[:2]: k for k, p in enumerate(pos)}
    def mask(cells):
        return sum(1 << idx[c] for c in cells)
    full = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    assert oracle.one_piece(mask(full), 8)
    assert not oracle.one_piece(mask([(-1, 0), (1, 0)]), 8)
    assert oracle.one_piece(mask([(1, 0)]), 8)
    assert oracle.one_piece(mask([(1, 0), (1, 1)]), 8)


def test_chemotaxis_work_reads_the_two_sites(oracle_lib):
    # chemotaxis.rs:109-119
    m = Model(width=16, height=16, contact=[0.0, 16.0, 16.0, 8.0],
              types=[CellType(), CellType(target_volume=64.0, lambda_volume=2.0)],
              species=[Species("attractant", 0.0)], secretion=[[0.0], [0.0]],
              uptake=[[0.0], [0.0]], chemotaxis=[[0.0], [2.0]], seed=4)
    labels = tiled_labels(m, 8, 2, 2, 1)
    sim = oracle.SerialSimulation(m, labels, default_cell_types(5))
    f = np.zeros(256)
    f[10], f[11] = 5.0, 1.0
    sim.set_field(0, f)
    assert abs(sim.chemotaxis_work(1, 10, 11) + 8.0) < 1e-12
    assert abs(sim.chemotaxis_work(1, 11, 10) - 8.0) < 1e-12
    assert sim.chemotaxis_work(0, 10, 11) == 0.0


def pop_model(division_volume, death_rate, target_volume, seed=7):
    # population.rs:7-36
    return Model(width=64, height=64, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), CellType(target_volume=target_volume, lambda_volume=2.0,
                                             division_volume=division_volume,
                                             death_rate=death_rate)],
                 seed=seed)


def test_a_division_splits_the_volume(oracle_lib):
    # population.rs:59-75
    sim = tiled(pop_model(40.0, 0.0, 100.0), side=8, nx=2, ny=2)
    before = sim.volumes()[1:].sum()
    divided, died = sim.divide_and_die()
    assert divided > 0 and died == 0
    assert sim.volumes()[1:].sum() == before
    assert (sim.volumes() == sim.recounted_volumes()).all()


def test_a_dying_cell_leaves_medium(oracle_lib):
    # population.rs:78-90
    sim = tiled(pop_model(0.0, 1.0, 64.0), side=8, nx=2, ny=2)
    _, died = sim.divide_and_die()
    assert died == 4
    assert (sim.labels() == 0).all()


def test_growing_cells_divide(oracle_lib):
    # population.rs:38-57
    sim = tiled(pop_model(72.0, 0.0, 100.0), side=6, nx=4, ny=4)
    sim.step(400)
    live = (sim.volumes()[1:] > 0).sum()
    assert live > 16
    assert (sim.volumes() == sim.recounted_volumes()).all()
    assert (sim.surfaces() == sim.recounted_surfaces()).all()


def test_death_takes_the_stated_fraction(oracle_lib):
    # population.rs:104-115
    sim = tiled(pop_model(0.0, 0.02, 64.0), side=8, nx=6, ny=6)
    sim.step(50)
    left = (sim.volumes()[1:] > 0).sum()
    expected = 36 * 0.98 ** 50
    assert abs(left - expected) < 0.35 * 36

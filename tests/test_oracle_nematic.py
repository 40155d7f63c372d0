"""The nematic coupling in the oracles, against the reference's own checks.

glazier-tests/tests/nematic.rs restated on the serial oracle (the f64
restatement of cpu.rs:315-332 and moments.rs:131-141): the incremental delta
against a full recomputation of the energy, the kept sums against a recount,
alignment with a uniform field and none without the coupling, and a labelled
restart. The emulation's term (the field rounded to multiples of 2^-24 and
summed per label in int64, the anisotropy in f64 from integer moments) is
checked against a from-scratch computation: the serial engine with the term
off, plus -lambda (F' . a' - F . a) with a from numpy on the cells' sites.
"""

import copy

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model
from tests.cases import tiled

SIDE = 48


def field_along(angle, amplitude=1.0, sites=SIDE * SIDE):
    # nematic.rs:9-17
    return np.tile([amplitude * np.cos(2 * angle), amplitude * np.sin(2 * angle)], (sites, 1))


def random_field(seed, sites=SIDE * SIDE):
    # nematic.rs:19-24: xoshiro256++ pairs next_f64() - 0.5
    rng = oracle.Xoshiro(seed)
    return np.array([[rng.next_f64() - 0.5, rng.next_f64() - 0.5] for _ in range(sites)])


def model(lambda_nematic, nematic, seed=5, side=SIDE):
    # nematic.rs:26-46
    c = CellType(target_volume=64.0, lambda_volume=2.0, lambda_nematic=lambda_nematic)
    return Model(width=side, height=side, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                 temperature=10.0, neighbour_order=2, seed=seed, nematic=nematic)


def order(sim):
    """nematic.rs:48-67: |<exp(2 i theta)>| of the long axes and the mean
    director, over cells of at least 8 sites, from recounted moments."""
    aniso = sim.recounted_anisotropy()
    vol = sim.recounted_volumes()
    c = s = n = 0.0
    for label in range(1, len(vol)):
        if vol[label] < 8:
            continue
        a0, a1 = aniso[label]
        norm = np.hypot(a0, a1)
        if norm < 1e-9:
            continue
        c += a0 / norm
        s += a1 / norm
        n += 1.0
    return np.hypot(c, s) / n, 0.5 * np.arctan2(s, c)


def test_the_incremental_delta_matches_a_full_recomputation(oracle_lib):
    # nematic.rs:69-113
    m = model(3.0, random_field(9))
    sim = oracle.SerialSimulation(m, *tiled(m))
    rng = oracle.Xoshiro(4)
    checked = 0
    for _ in range(20):
        sim.step()
        for _ in range(40):
            lab = sim.labels()
            target = rng.below(lab.size)
            source = rng.below(lab.size)
            new, old = int(lab[source]), int(lab[target])
            if old == new or (old != 0 and sim.volumes()[old] <= 1):
                continue
            predicted = sim.delta_energy(target, new)
            before = sim.energy()
            after = sim.clone()
            after.apply_copy(target, new)
            measured = after.energy() - before
            after.close()
            assert abs(predicted - measured) < 1e-8, (target, predicted, measured)
            checked += 1
    assert checked > 100


def test_the_nematic_sums_match_a_recount_after_a_run(oracle_lib):
    # nematic.rs:115-128
    m = model(3.0, random_field(1))
    sim = oracle.SerialSimulation(m, *tiled(m))
    sim.step(60)
    assert np.allclose(sim.nematic_sums(), sim.recounted_nematic(), rtol=0.0, atol=1e-8)


def test_cells_align_with_a_uniform_field_and_not_without_one(oracle_lib):
    # nematic.rs:130-148
    angle = 0.6
    coupled = oracle.SerialSimulation(model(6.0, field_along(angle)), *tiled(model(6.0, None)))
    blind = oracle.SerialSimulation(model(0.0, field_along(angle)), *tiled(model(0.0, None)))
    coupled.step(400)
    blind.step(400)
    s_coupled, director = order(coupled)
    s_blind, _ = order(blind)
    off = min(abs(director - angle), np.pi - abs(director - angle))
    print(f"[nematic] serial order {s_coupled:.3f} director {director:.3f}, uncoupled {s_blind:.3f}")
    assert s_coupled > 0.7
    assert off < 0.15
    assert s_blind < 0.4


def test_a_labelled_start_reproduces_the_state_it_was_taken_from(oracle_lib):
    # nematic.rs:150-166
    m = model(2.0, random_field(3))
    first = oracle.SerialSimulation(m, *tiled(m))
    first.step(30)
    restarted = oracle.SerialSimulation(m, first.labels(), first.cell_types())
    assert np.array_equal(restarted.volumes(), first.recounted_volumes())
    assert np.array_equal(restarted.surfaces(), first.recounted_surfaces())
    assert np.allclose(restarted.lengths()[1:], first.lengths()[1:], rtol=0.0, atol=1e-9)
    assert np.allclose(restarted.nematic_sums()[1:], first.nematic_sums()[1:], rtol=0.0, atol=1e-9)
    assert abs(restarted.energy() - first.energy()) < 1e-6


def test_a_field_of_the_wrong_size_is_refused():
    # nematic.rs:168-172 (model.rs:193-197)
    m = model(0.0, np.zeros((10, 2)))
    with pytest.raises(ValueError, match="sites"):
        m.validate()


def aniso_from_sites(labels, m, label):
    """(a0, a1) of one label from its sites, unwrapped about its first site by
    minimum image, population covariance in f64."""
    w, h = m.width, m.height
    idx = np.nonzero(labels == label)[0]
    if idx.size < 2:
        return 0.0, 0.0
    xy = np.stack([idx % w, (idx // w) % h], axis=1).astype(np.float64)
    d = xy - xy[0]
    for ax, span in enumerate((w, h)):
        d[:, ax] -= span * np.round(d[:, ax] / span)
    c = np.cov(d.T, bias=True)
    tr = c[0, 0] + c[1, 1]
    if tr <= 1e-12:
        return 0.0, 0.0
    return (c[0, 0] - c[1, 1]) / tr, 2.0 * c[0, 1] / tr


def test_emulated_nematic_changes_match_a_from_scratch_computation(oracle_lib):
    m = model(3.0, random_field(9, 64 * 64), seed=11, side=64)
    emu = oracle.EmuSimulation(m, *tiled(m))
    emu.step(4)
    lab = emu.labels()
    m0 = copy.deepcopy(m)
    m0.types[1].lambda_nematic = 0.0
    ser = oracle.SerialSimulation(m0, lab, emu.cell_types())
    q = np.array([[oracle_lib.emu_nematic_fixed(v) for v in row] for row in m.nematic],
                 dtype=np.float64) * 2.0 ** -24
    lam = m.types[1].lambda_nematic
    rng = np.random.default_rng(3)
    w = m.width
    checked = 0
    for _ in range(600):
        t = int(rng.integers(lab.size))
        x, y = t % w, t // w
        dx, dy = rng.integers(-1, 2, size=2)
        s_ = int(((y + dy) % m.height) * w + (x + dx) % w)
        old, new = int(lab[t]), int(lab[s_])
        if old == new:
            continue
        after = lab.copy()
        after[t] = new
        want = ser.delta_energy(t, new)
        for c in (old, new):
            if c == 0:
                continue
            f0 = q[lab == c].sum(axis=0)
            f1 = q[after == c].sum(axis=0)
            a0 = aniso_from_sites(lab, m, c)
            a1 = aniso_from_sites(after, m, c)
            want += -lam * ((f1[0] * a1[0] + f1[1] * a1[1]) - (f0[0] * a0[0] + f0[1] * a0[1]))
        got = emu.delta_energy(t, s_)
        assert abs(got - want) <= 2e-5 * max(1.0, abs(want)), (t, s_, got, want)
        checked += 1
    assert checked > 50


def test_emulated_nematic_sums_are_exact_fixed_point_recounts(oracle_lib):
    m = model(3.0, random_field(5, 64 * 64), seed=12, side=64)
    emu = oracle.EmuSimulation(m, *tiled(m))
    emu.step(10)
    lab = emu.labels()
    state = emu.cell_state()
    assert state.shape[1] == 13
    qf = np.array([[oracle_lib.emu_nematic_fixed(v) for v in row] for row in m.nematic],
                  dtype=np.int64)
    want = np.zeros((state.shape[0], 2), dtype=np.int64)
    np.add.at(want, lab, qf)
    assert np.array_equal(state[1:, 11:13], want[1:])


def test_the_emulated_schedule_aligns_cells_as_the_serial_engine_does(oracle_lib):
    # gpu_nematic.rs:70-104 on the emulated device schedule
    angle = 0.6
    m = model(6.0, field_along(angle, sites=64 * 64), seed=7, side=64)
    emu = oracle.EmuSimulation(m, *tiled(m))
    emu.step(400)
    check = oracle.SerialSimulation(m, emu.labels(), emu.cell_types())
    s_emu, director = order(check)
    off = min(abs(director - angle), np.pi - abs(director - angle))
    print(f"[nematic] emulated order {s_emu:.3f} director {director:.3f}")
    assert s_emu > 0.7 and off < 0.15

"""The HIP engine against the serial f64 engine, in distribution.

The checks of glazier-tests/tests/gpu_matches_cpu.rs, one for one, with the
serial oracle (the restatement of glazier-core's engine) in the serial
engine's place; the first runs over 32 seeds instead of one. The two
engines sample the same model by different chains, so these compare what a
run is read for: cell counts, volumes and their spread about target, surface
and length bookkeeping, fields, drift up a gradient. At the 4096^2 scale the
checks are the size-independent ones: conservation, the device counters
against a recount of the lattice they wrote, and determinism.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import Species
from tests.cases import lone_cell, ref_model, tiled

pytestmark = pytest.mark.gpu


def spread(v, target):
    return float(np.sqrt(((np.asarray(v, dtype=np.float64) - target) ** 2).mean()))


def gpu_run(m, labels, types, steps, fields=None):
    from glazier_amd import HipSimulation

    g = HipSimulation(m, labels, types, fields)
    g.step(steps)
    return g


def test_the_gpu_reproduces_the_serial_run_in_distribution_over_32_seeds(engine, oracle_lib):
    # gpu_matches_cpu.rs:48-92
    sc, sg = [], []
    for seed in range(1, 33):
        m = ref_model(64, 64, seed=seed)
        l, t = tiled(m)
        cpu = oracle.SerialSimulation(m, l, t)
        cpu.step(200)
        gpu = gpu_run(m, l, t, 200)
        vc, vg_all = cpu.volumes()[1:], gpu.volumes()
        vg = vg_all[1:]
        assert len(vc) == len(vg), "cell counts differ"
        assert abs(vc.mean() - vg.mean()) < 1e-9
        recounted = np.bincount(gpu.labels(), minlength=vg_all.size)
        assert np.array_equal(recounted, vg_all), "the device volume counter drifted"
        sc.append(spread(vc, 64.0))
        sg.append(spread(vg, 64.0))
    a, b = np.mean(sc), np.mean(sg)
    print(f"[distribution] spread about target: serial {a:.3f}, device {b:.3f}", flush=True)
    assert abs(a - b) / max(a, b) < 0.25, (a, b)


def test_the_device_surface_counter_matches_the_lattice_it_wrote(engine, oracle_lib):
    # gpu_matches_cpu.rs:95-119
    m = ref_model(64, 64, seed=3)
    m.types[1].target_surface = 34.0
    m.types[1].lambda_surface = 0.1
    l, t = tiled(m)
    gpu = gpu_run(m, l, t, 150)
    check = oracle.SerialSimulation(m, l, t)
    check.set_labels(gpu.labels())
    assert np.array_equal(gpu.surfaces(), check.recounted_surfaces())


def test_a_device_run_conserves_its_cells(engine, oracle_lib):
    # gpu_matches_cpu.rs:121-138
    m = ref_model(64, 64, seed=7)
    l, t = tiled(m)
    gpu = gpu_run(m, l, t, 300)
    v = gpu.volumes()
    assert v.size - 1 == int(l.max())
    assert (v[1:] >= 1).all()


def test_the_device_stretches_cells_the_way_the_serial_engine_does(engine, oracle_lib):
    # gpu_matches_cpu.rs:156-204
    stated = ref_model(64, 64, seed=41, target_length=24.0, lambda_length=1.0)
    l, t = tiled(stated)
    cpu = oracle.SerialSimulation(stated, l, t)
    cpu.step(200)
    gpu = gpu_run(stated, l, t, 200)
    on_device = oracle.SerialSimulation(stated, l, t)
    on_device.set_labels(gpu.labels())
    loose_m = ref_model(64, 64, seed=41)
    loose = oracle.SerialSimulation(loose_m, l, t)
    loose.step(200)
    serial = cpu.recounted_lengths()[1:].mean()
    device = on_device.recounted_lengths()[1:].mean()
    free = loose.recounted_lengths()[1:].mean()
    print(f"[length] serial {serial:.3f} device {device:.3f} free {free:.3f}", flush=True)
    assert serial > free and device > free
    assert abs(serial - device) / max(serial, device) < 0.15


def secreting(seed):
    m = ref_model(64, 64, seed=seed)
    m.species = [Species("IFNg", 0.4, 0.01, 0.0)]
    m.secretion = [[0.0], [0.2]]
    m.uptake = [[0.0], [0.01]]
    m.chemotaxis = [[0.0], [0.0]]
    return m


def test_the_device_solves_a_field_the_way_the_serial_engine_does(engine, oracle_lib):
    # gpu_matches_cpu.rs:227-255
    m = secreting(11)
    l, t = tiled(m)
    cpu = oracle.SerialSimulation(m, l, t)
    cpu.step(60)
    gpu = gpu_run(m, l, t, 60)
    device, serial = gpu.field(0).sum(), cpu.field(0).sum()
    assert device > 0.0
    assert abs(device - serial) / max(serial, device) < 0.02


def test_a_device_field_stays_where_it_started_without_sources(engine, oracle_lib):
    # gpu_matches_cpu.rs:257-280
    m = secreting(12)
    m.secretion = [[0.0], [0.0]]
    m.uptake = [[0.0], [0.0]]
    m.species[0].decay = 0.0
    m.species[0].initial = 2.0
    l, t = tiled(m)
    gpu = gpu_run(m, l, t, 40)
    sites = m.width * m.height
    assert abs(gpu.field(0).sum() - 2.0 * sites) / (2.0 * sites) < 1e-4


def test_the_device_climbs_a_gradient_the_way_the_serial_engine_does(engine, oracle_lib):
    # gpu_matches_cpu.rs:282-330
    m = ref_model(64, 64, seed=21)
    m.species = [Species("attractant", 0.0, 0.0, 0.0)]
    m.secretion = [[0.0], [0.0]]
    m.uptake = [[0.0], [0.0]]
    m.chemotaxis = [[0.0], [4.0]]
    l, t = lone_cell(m, 24, 32)
    ramp = np.tile(np.arange(64, dtype=np.float64), 64)

    def centre(labels):
        xs = np.nonzero(labels == 1)[0] % 64
        return xs.mean()

    start = centre(l)
    gpu = gpu_run(m, l, t, 300, fields=ramp)
    moved = centre(gpu.labels()) - start
    assert moved > 3.0, moved


def test_the_device_runs_a_volume(engine, oracle_lib):
    # gpu_matches_cpu.rs:332-384
    m = ref_model(32, 32, 32, seed=31)
    m.types[1].target_volume = 512.0
    l, t = tiled(m)
    cpu = oracle.SerialSimulation(m, l, t)
    cpu.step(40)
    gpu = gpu_run(m, l, t, 40)
    dv, cv = gpu.volumes(), cpu.volumes()
    assert dv.size == cv.size
    assert dv[1:].sum() == cv[1:].sum()
    counted = np.bincount(gpu.labels(), minlength=dv.size)
    assert np.array_equal(dv[1:], counted[1:])
    sc, sg = spread(cv[1:], 512.0), spread(dv[1:], 512.0)
    assert abs(sc - sg) / max(sc, sg) < 0.3, (sc, sg)


def surfaces_of(labels, w, h):
    """Unlike Moore bonds per label (count_surface, glazier-core/src/cpu.rs:11-59)."""
    lab = labels.reshape(h, w)
    n = int(labels.max()) + 1
    out = np.zeros(n, dtype=np.int64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx == 0 and dy == 0:
                continue
            other = np.roll(np.roll(lab, -dy, axis=0), -dx, axis=1)
            out += np.bincount(lab[lab != other], minlength=n)
    return out


def test_a_4096_lattice_keeps_its_books(engine, oracle_lib):
    """At the measured size: the sheet stays confluent, every counter matches
    a recount of the lattice, and a second run of the same seed is the same."""
    from glazier_amd.workloads import plain_monolayer

    m, l, t = plain_monolayer(4096, seed=2, surface=False)
    a = gpu_run(m, l, t, 6)
    labels = a.labels()
    vol = a.volumes()
    assert vol.size == 262145 and (vol[1:] >= 1).all()
    assert np.array_equal(np.bincount(labels, minlength=vol.size), vol)
    assert vol[1:].sum() == 4096 * 4096
    surf = a.surfaces()
    assert np.array_equal(surf[1:], surfaces_of(labels, 4096, 4096)[1:])
    attempted, accepted = a.counters()
    assert attempted == 6 * 4096 * 4096 and accepted > 0
    a.close()
    b = gpu_run(m, l, t, 6)
    assert np.array_equal(b.labels(), labels)


def test_the_4096_sweep_matches_the_serial_engine_in_distribution(engine, oracle_lib):
    """The measured size against the f64 serial engine, in distribution: the
    reference's benchmark model (examples/bench.rs:7-25: contact 16/8, V_t 64,
    lambda_V 2, T 10, Moore), 100 steps. Per-cell statistics do not depend on
    the lattice's size, so the serial side runs four 256^2 sheets (seeds
    1-4, ~2 s) against one 4096^2 device sheet of 262144 cells. Criteria
    as gpu_matches_cpu.rs:48-92: every cell kept, the same mean volume, the
    volume spread about target within a quarter; and the mean surface within
    5%."""
    from glazier_amd.model import CellType, Model, default_cell_types, tiled_labels

    def bench_model(side, seed):
        c = CellType(target_volume=64.0, lambda_volume=2.0)
        m = Model(width=side, height=side, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                  temperature=10.0, neighbour_order=2, seed=seed)
        n = side // 8
        return m, tiled_labels(m, 8, n, n, 1), default_cell_types(n * n + 1)

    sc, surf_c = [], []
    for seed in (1, 2, 3, 4):
        cpu = oracle.SerialSimulation(*bench_model(256, seed))
        cpu.step(100)
        v = cpu.volumes()[1:]
        assert v.size == 1024 and v.mean() == 64.0
        sc.append(spread(v, 64.0))
        surf_c.append(cpu.surfaces()[1:].mean())
    gpu = gpu_run(*bench_model(4096, 1), 100)
    vg = gpu.volumes()[1:]
    assert vg.size == 262144 and (vg >= 1).all() and vg.mean() == 64.0
    sg = spread(vg, 64.0)
    surf_g = gpu.surfaces()[1:].mean()
    a, b = float(np.mean(sc)), sg
    print(f"[distribution 4096] spread about target: serial {a:.3f}, device {b:.3f}; "
          f"mean surface: serial {np.mean(surf_c):.2f}, device {surf_g:.2f}", flush=True)
    assert abs(a - b) / max(a, b) < 0.25, (a, b)
    assert abs(np.mean(surf_c) - surf_g) / np.mean(surf_c) < 0.05
    gpu.close()

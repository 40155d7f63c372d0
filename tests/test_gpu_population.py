"""Division and death on the device, against the serial oracle.

glazier-tests/tests/gpu_population.rs restated, one for one: the death rate
on the device, a dead cell leaving medium, division conserving sites, the
device counters against the lattice they wrote, and both engines losing cells
at the rate the description states.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model
from tests.cases import tiled

pytestmark = pytest.mark.gpu


def model(division_volume, death_rate, target_volume, seed):
    # gpu_population.rs:9-29
    c = CellType(target_volume=target_volume, lambda_volume=2.0, division_volume=division_volume,
                 death_rate=death_rate)
    return Model(width=64, height=64, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                 temperature=10.0, neighbour_order=2, seed=seed)


def live(volumes):
    return int((np.asarray(volumes)[1:] > 0).sum())


def device(m, side, nx, ny):
    from glazier_amd import HipSimulation

    labels, types = tiled(m, side=side, nx=nx, ny=ny)
    return HipSimulation(m, labels, types), int(labels.max())


def test_the_device_kills_cells_at_the_stated_rate(engine):
    # gpu_population.rs:45-60
    gpu, started = device(model(0.0, 0.02, 64.0, 5), 8, 6, 6)
    gpu.step(50)
    left = live(gpu.volumes())
    expected = started * 0.98 ** 50
    assert abs(left - expected) < 0.35 * started, (left, expected, started)
    gpu.close()


def test_a_dead_cell_leaves_medium_on_the_device(engine):
    # gpu_population.rs:62-79
    gpu, _ = device(model(0.0, 1.0, 64.0, 6), 8, 2, 2)
    gpu.step(1)
    assert live(gpu.volumes()) == 0, "a death rate of one takes every cell"
    assert (gpu.labels() == 0).all(), "a dead cell left sites behind"
    gpu.close()


def test_the_device_divides_and_conserves_the_sites(engine):
    # gpu_population.rs:81-102
    gpu, started = device(model(40.0, 0.0, 100.0, 7), 8, 4, 4)
    gpu.step(1)
    volumes = gpu.volumes()
    assert live(volumes) > started, (live(volumes), started)
    assert int(volumes[1:].sum()) == int((gpu.labels() != 0).sum()), \
        "division moved sites rather than making them"
    gpu.close()


def test_the_device_counters_match_the_lattice_they_wrote(engine):
    # gpu_population.rs:104-119
    gpu, _ = device(model(72.0, 0.005, 100.0, 8), 6, 6, 6)
    gpu.step(120)
    volumes = gpu.volumes()
    counted = np.bincount(gpu.labels(), minlength=volumes.size)
    assert np.array_equal(volumes[1:], counted[1:volumes.size])
    gpu.close()


def test_both_engines_lose_cells_at_the_rate_the_description_states(engine, oracle_lib):
    # gpu_population.rs:121-157: each within three binomial spreads of the
    # expected survivors
    steps, rate = 80, 0.01
    m = model(0.0, rate, 64.0, 9)
    labels, types = tiled(m, side=8, nx=10, ny=10)
    started = int(labels.max())
    cpu = oracle.SerialSimulation(m, labels, types)
    cpu.step(steps)
    gpu, _ = device(m, 8, 10, 10)
    gpu.step(steps)
    survival = (1.0 - rate) ** steps
    expected = started * survival
    spread = np.sqrt(started * survival * (1.0 - survival))
    for name, left in (("serial", live(cpu.volumes())), ("device", live(gpu.volumes()))):
        assert abs(left - expected) < 3.0 * spread, (name, left, expected, spread)
    gpu.close()


# ---- label capacity: grown as the reference grows its per-label buffers
# (glazier-cuda/src/lib.rs:1446-1459), so divisions past the first
# allocation are never dropped and the run stays the emulation's.

def dividing_sheet(seed):
    # 4x4 cells that want 32 sites and divide at 8: every cell divides in the
    # first step and about half again in the second, past the 2n + 1024 ids
    # set aside at creation (n = 4097 labels, 9218 ids)
    c = CellType(target_volume=32.0, lambda_volume=2.0, division_volume=8.0)
    m = Model(width=256, height=256, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
              temperature=10.0, neighbour_order=2, seed=seed)
    labels, types = tiled(m, side=4, nx=64, ny=64)
    return m, labels, types


def test_divisions_past_the_first_allocation_match_the_emulation(engine):
    from glazier_amd import HipSimulation

    m, labels, types = dividing_sheet(11)
    gpu = HipSimulation(m, labels, types)
    emu = oracle.EmuSimulation(m, labels, types)
    for chunk in (1, 5):  # a call that divides at once, then one that runs on
        gpu.step(chunk)
        emu.step(chunk)
        assert np.array_equal(gpu.labels(), emu.labels())
        assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.n_labels() > 2 * 4097 + 1024, gpu.n_labels()
    gpu.close()


def test_a_batch_grows_its_lattices_and_matches_the_emulation(engine):
    from glazier_amd import HipBatch, HipSimulation

    sims, emus = [], []
    for seed in (12, 13):
        m, labels, types = dividing_sheet(seed)
        sims.append(HipSimulation(m, labels, types))
        emus.append(oracle.EmuSimulation(m, labels, types))
    with HipBatch(sims) as batch:
        batch.set_tiles(0)
        batch.step(4)
    for gpu, emu in zip(sims, emus):
        emu.step(4)
        assert gpu.n_labels() > 2 * 4097 + 1024
        assert np.array_equal(gpu.labels(), emu.labels())
        assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
        gpu.close()

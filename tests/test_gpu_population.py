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

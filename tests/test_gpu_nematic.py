"""The nematic coupling on the device.

glazier-tests/tests/gpu_nematic.rs restated: cells line up along a uniform
field by about as much as the serial engine lines them up, not at all when
the coupling is zero, and turn with a replaced field; a field of the wrong
size is refused. Bit-exact parity of the term with the emulation is in
test_gpu_parity.py (cases nematic, nematic_division, nematic3d) and below for
a field replaced mid-run.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model
from tests.cases import case, random_pairs, tiled

pytestmark = pytest.mark.gpu

SIDE = 64


def model(lambda_nematic, angle):
    # gpu_nematic.rs:15-37
    c = CellType(target_volume=64.0, lambda_volume=2.0, lambda_nematic=lambda_nematic)
    field = np.tile([np.cos(2 * angle), np.sin(2 * angle)], (SIDE * SIDE, 1))
    return Model(width=SIDE, height=SIDE, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                 temperature=10.0, neighbour_order=2, seed=7, nematic=field)


def order(m, labels, types):
    """gpu_nematic.rs:39-61: order and director of a label field's cells, from
    moments recounted by the serial oracle."""
    sim = oracle.SerialSimulation(m, labels, types)
    aniso = sim.recounted_anisotropy()
    vol = sim.recounted_volumes()
    sim.close()
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


def off(d, a):
    return min(abs(d - a), np.pi - abs(d - a))


def run_gpu(m, steps):
    from glazier_amd import HipSimulation

    labels, types = tiled(m)
    g = HipSimulation(m, labels, types)
    g.step(steps)
    out = g.labels(), g.cell_types()
    g.close()
    return out


def test_the_device_aligns_cells_along_the_field_as_the_serial_engine_does(engine, oracle_lib):
    # gpu_nematic.rs:70-104
    angle, steps = 0.6, 400
    coupled = model(6.0, angle)
    labels, types = run_gpu(coupled, steps)
    s_gpu, director = order(coupled, labels, types)
    cpu = oracle.SerialSimulation(coupled, *tiled(coupled))
    cpu.step(steps)
    s_cpu, _ = order(coupled, cpu.labels(), cpu.cell_types())
    blind = model(0.0, angle)
    s_blind, _ = order(blind, *run_gpu(blind, steps))
    print(f"[nematic] device order {s_gpu:.3f} director {director:.3f}, serial {s_cpu:.3f}, "
          f"uncoupled device {s_blind:.3f}", flush=True)
    assert s_gpu > 0.7
    assert off(director, angle) < 0.15
    assert abs(s_gpu - s_cpu) < 0.15
    assert s_blind < 0.4


def test_a_replaced_field_turns_the_cells_on_the_device(engine, oracle_lib):
    # gpu_nematic.rs:106-136
    from glazier_amd import GhipError, HipSimulation

    first = 0.2
    second = first + np.pi / 2
    coupled = model(6.0, first)
    labels, types = tiled(coupled)
    gpu = HipSimulation(coupled, labels, types)
    gpu.step(400)
    _, before = order(coupled, gpu.labels(), gpu.cell_types())
    turned = np.tile([np.cos(2 * second), np.sin(2 * second)], (SIDE * SIDE, 1))
    gpu.set_nematic_field(turned)
    gpu.step(800)
    s_after, after = order(coupled, gpu.labels(), gpu.cell_types())
    print(f"[nematic] director before {before:.3f}, after {after:.3f} (order {s_after:.3f})", flush=True)
    assert off(before, first) < 0.15
    assert s_after > 0.7
    assert off(after, second) < 0.15
    with pytest.raises(GhipError, match="sites"):
        gpu.set_nematic_field(turned[1:])
    gpu.close()


def test_a_replaced_field_matches_the_emulation_bit_for_bit(engine, oracle_lib):
    from glazier_amd import HipSimulation

    m, labels, types, _, _, _ = case("nematic")
    print("[parity] nematic field replaced after 7 of 16 steps", flush=True)
    gpu = HipSimulation(m, labels, types)
    emu = oracle.EmuSimulation(m, labels, types)
    gpu.step(7)
    emu.step(7)
    q = random_pairs(11, m.sites)
    gpu.set_nematic_field(q)
    emu.set_nematic_field(q)
    gpu.step(9)
    emu.step(9)
    assert np.array_equal(gpu.labels(), emu.labels())
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.counters() == emu.counters()
    gpu.close()


def test_a_model_without_the_coupling_refuses_a_field(engine):
    from glazier_amd import GhipError, HipSimulation

    m = model(0.0, 0.3)  # a field, but no type couples to it
    labels, types = tiled(m)
    gpu = HipSimulation(m, labels, types)
    with pytest.raises(GhipError, match="no nematic coupling"):
        gpu.set_nematic_field(m.nematic)
    gpu.close()

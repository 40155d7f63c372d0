"""Batches of lattices against the CPU emulation, bit for bit.

A lattice stepped in a batch must end exactly where the emulation of its own
schedule ends: the batch only changes how launches are grouped (and which
tile shape a pass uses), never a draw or an acceptance. The cases mix planes
and volumes, every priced term, division/death, and the tile candidates the
batch times on its first steps.
"""

import numpy as np
import pytest

import oracle
from tests.cases import case

pytestmark = pytest.mark.gpu

BATCH_CASES = ["plain128", "fast_medium_typed", "medium_gaps", "rect64x32", "von_neumann",
               "float_energies", "volume3d_o2", "volume3d_o3", "volume3d_o1", "connected",
               "connected3d", "chemotaxis_ramp", "secreting", "length", "motility", "drift",
               "death", "division", "infection128", "infection3d", "motility3d", "length3d", "division3d",
               "nematic", "nematic_division", "nematic3d"]


def assert_same(name, m, gpu, emu):
    assert gpu.n_labels() == emu.n_labels(), name
    diff = int((gpu.labels() != emu.labels()).sum())
    assert diff == 0, f"{name}: {diff} sites differ"
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:]), name
    assert np.array_equal(gpu.volumes(), emu.volumes()), name
    assert np.array_equal(gpu.surfaces(), emu.surfaces()), name
    assert np.array_equal(gpu.cell_types(), emu.cell_types()), name
    assert gpu.counters() == emu.counters(), name
    for k in range(len(m.species)):
        assert np.array_equal(gpu.field_f32(k), emu.field_f32(k)), f"{name}: species {k}"
    assert np.array_equal(gpu.activity_f32(), emu.activity_f32()), name


def build(names):
    from glazier_amd import HipSimulation

    runs = []
    for name in names:
        m, labels, types, fields, activity, _ = case(name)
        runs.append((name, m, HipSimulation(m, labels, types, fields, activity),
                     oracle.EmuSimulation(m, labels, types, fields, activity)))
    return runs


def test_a_batch_matches_each_lattice_alone(engine, oracle_lib):
    from glazier_amd import HipBatch

    runs = build(BATCH_CASES)
    print(f"[batch] {len(runs)} lattices, 2 + 12 steps", flush=True)
    batch = HipBatch([r[2] for r in runs])
    batch.step(2)   # warm-up and the first timed tile candidate
    batch.step(12)  # the other tiles, the occupancy caps, then the captured step
    info = batch.describe()
    assert info["tuned"] and info["graph_steps"] == 5 and info["eager_steps"] == 9
    assert {f["dims"] for f in info["fields"]} == {2, 3}
    for g in info["groups"] + info["fields"]:
        assert g["wgs_per_cu"] in info["wgs_per_cu_candidates"]
        assert all(ms > 0 for ms in g["tuning_ms"] + g["occupancy_ms"])
    for name, m, gpu, emu in runs:
        emu.step(14)
        assert gpu.mcs == 14
        assert_same(name, m, gpu, emu)
    batch.close()


@pytest.mark.parametrize("candidate", range(5))
def test_every_tile_candidate_gives_the_same_lattice(engine, oracle_lib, candidate):
    from glazier_amd import HipBatch

    runs = build(["surface256", "length", "volume3d_o2", "division", "motility3d", "infection128",
                  "infection3d"])
    print(f"[batch] tile candidate {candidate}", flush=True)
    batch = HipBatch([r[2] for r in runs])
    batch.set_tiles(candidate)
    batch.step(4)
    assert batch.describe()["graph_steps"] == 4
    for name, m, gpu, emu in runs:
        emu.step(4)
        assert_same(name, m, gpu, emu)
    batch.close()


@pytest.mark.parametrize("candidate", range(1, 4))
def test_every_occupancy_cap_gives_the_same_lattice(engine, oracle_lib, candidate):
    from glazier_amd import HipBatch

    runs = build(["infection128", "infection3d", "volume3d_o2", "chemotaxis_ramp", "division3d"])
    print(f"[batch] occupancy candidate {candidate}", flush=True)
    batch = HipBatch([r[2] for r in runs])
    batch.set_tiles(1)
    batch.set_occupancy(candidate)
    batch.step(3)
    info = batch.describe()
    assert info["graph_steps"] == 3
    want = info["wgs_per_cu_candidates"][candidate]
    assert all(g["wgs_per_cu"] == want for g in info["groups"] + info["fields"])
    for name, m, gpu, emu in runs:
        emu.step(3)
        assert_same(name, m, gpu, emu)
    batch.close()


def test_a_lattice_stepped_alone_between_batch_steps(engine, oracle_lib):
    from glazier_amd import HipBatch

    # alone, plain128 steps on the resident single-workgroup pass and
    # surface256 on the tiled resident pass (one workgroup per tile, grid
    # barriers), between steps of the batch's own passes over the same buffers
    runs = build(["infection128", "motility", "connected3d", "plain128", "surface256"])
    batch = HipBatch([r[2] for r in runs])
    batch.step(7)
    for i in (1, 3, 4):
        runs[i][2].step(2)
    batch.step(3)
    for i, (name, m, gpu, emu) in enumerate(runs):
        emu.step(12 if i in (1, 3, 4) else 10)
        assert_same(name, m, gpu, emu)
    batch.close()


def test_a_small_uneven_ensemble(engine, oracle_lib):
    from glazier_amd import HipBatch, HipSimulation
    from glazier_amd.workloads import ensemble

    members = ensemble(seed=5, planar=6, volumes=2, sides=(64, 128), volume_side=16)
    sims = [HipSimulation(m, l, t) for m, l, t in members]
    batch = HipBatch(sims)
    batch.step(8)
    for (m, l, t), gpu in zip(members, sims):
        emu = oracle.EmuSimulation(m, l, t)
        emu.step(8)
        assert_same(f"replica {m.replica}", m, gpu, emu)
    batch.close()


def test_a_lattice_in_a_batch_is_not_destroyed(engine, oracle_lib):
    from glazier_amd import GhipError, HipBatch

    runs = build(["ref64", "plain128"])
    batch = HipBatch([r[2] for r in runs])
    with pytest.raises(GhipError):
        runs[0][2].close()
    with pytest.raises(GhipError):
        HipBatch([runs[1][2]])  # already in a batch
    batch.close()
    for r in runs:
        r[2].close()

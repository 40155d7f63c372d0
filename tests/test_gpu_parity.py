"""The HIP engine against the CPU emulation of its schedule, bit for bit.

Label lattices, per-label volume/surface/moment sums, acceptance counts, f32
fields and f32 activity must be identical after the same steps from the same
seed. The cases cover every term the sweep prices, planes and volumes, all
three neighbour orders, small and odd-shaped lattices, and division/death.
"""

import sys

import numpy as np
import pytest

import oracle
from tests.cases import PARITY_CASES, case

pytestmark = pytest.mark.gpu


def run_both(name):
    from glazier_amd import HipSimulation

    m, labels, types, fields, activity, steps = case(name)
    print(f"[parity] {name}: {m.width}x{m.height}x{m.depth}, {steps} MCS", flush=True)
    gpu = HipSimulation(m, labels, types, fields, activity)
    emu = oracle.EmuSimulation(m, labels, types, fields, activity)
    gpu.step(steps)
    emu.step(steps)
    return m, gpu, emu


@pytest.mark.parametrize("name", PARITY_CASES)
def test_lattice_and_bookkeeping_match_the_emulation(engine, oracle_lib, name):
    m, gpu, emu = run_both(name)
    assert gpu.n_labels() == emu.n_labels()
    got, want = gpu.labels(), emu.labels()
    diff = int((got != want).sum())
    assert diff == 0, f"{diff} sites differ"
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert np.array_equal(gpu.volumes(), emu.volumes())
    assert np.array_equal(gpu.surfaces(), emu.surfaces())
    assert np.array_equal(gpu.cell_types(), emu.cell_types())
    assert gpu.counters() == emu.counters()
    for k in range(len(m.species)):
        assert np.array_equal(gpu.field_f32(k), emu.field_f32(k)), f"species {k}"
    assert np.array_equal(gpu.activity_f32(), emu.activity_f32())
    gpu.close()


@pytest.mark.parametrize("order,w,h,scatter", [(1, 512, 256, False), (2, 256, 512, False), (2, 512, 256, True)])
def test_the_plane_recount_matches_the_emulation(engine, oracle_lib, order, w, h, scatter):
    # a plane whose records are volume and surface alone is recounted in
    # blocks of 16 rows by 256 columns: cells of side 7 cut by the blocks'
    # edges, a medium margin, both planar neighbourhoods, and (scattered
    # sites) more labels in a block than its table holds
    from glazier_amd import HipSimulation
    from tests.cases import ref_model, tiled

    m = ref_model(w, h, seed=41, order=order)
    labels, types = tiled(m, side=7)
    if scatter:
        labels = np.random.default_rng(41).permutation(labels)
    print(f"[parity] plane recount {w}x{h}, order {order}, scattered {scatter}", flush=True)
    gpu = HipSimulation(m, labels, types)
    emu = oracle.EmuSimulation(m, labels, types)
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert np.array_equal(gpu.volumes(), emu.volumes())
    assert np.array_equal(gpu.surfaces(), emu.surfaces())
    gpu.step(3)
    emu.step(3)
    assert int((gpu.labels() != emu.labels()).sum()) == 0
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.counters() == emu.counters()
    gpu.close()
    emu.close()


def test_a_second_call_continues_the_same_run(engine, oracle_lib):
    from glazier_amd import HipSimulation

    m, labels, types, fields, activity, _ = case("infection128")
    gpu = HipSimulation(m, labels, types)
    for chunk in (3, 1, 6):
        gpu.step(chunk)
    emu = oracle.EmuSimulation(m, labels, types)
    emu.step(10)
    assert np.array_equal(gpu.labels(), emu.labels())
    assert gpu.mcs == 10


def test_the_device_refuses_an_odd_lattice(engine):
    # gpu_matches_cpu.rs:140-149
    from glazier_amd import GhipError, HipSimulation
    from tests.cases import ref_model, tiled

    m = ref_model(33, 64)
    l, t = tiled(m)
    with pytest.raises(GhipError) as err:
        HipSimulation(m, l, t)
    assert "even sides" in str(err.value)


@pytest.mark.parametrize("name,chunks", [("ref64", (70, 3, 1)), ("fast_medium_typed", (65,)),
                                         ("rect64x32", (1, 64, 2))])
def test_a_resident_lattice_matches_across_launch_and_call_boundaries(engine, oracle_lib, name, chunks):
    # small plain lattices step inside one workgroup, at most 64 steps per
    # launch: runs that cross those launches and calls match the emulation
    from glazier_amd import HipSimulation

    m, labels, types, _, _, _ = case(name)
    steps = sum(chunks)
    print(f"[parity] resident {name}: {m.width}x{m.height}, chunks {chunks}", flush=True)
    gpu = HipSimulation(m, labels, types)
    gpu.set_timing(True)
    for c in chunks:
        gpu.step(c)
    _, launches = gpu.kernel_time("sweep")
    assert launches == sum((c + 63) // 64 for c in chunks), "the resident pass did not run"
    emu = oracle.EmuSimulation(m, labels, types)
    emu.step(steps)
    assert np.array_equal(gpu.labels(), emu.labels())
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.counters() == emu.counters()
    assert gpu.mcs == steps
    gpu.close()


def coop_case(name):
    from tests.cases import ref_model, tiled

    if name == "ref256":
        m = ref_model(256, 256, seed=21)
        return (m, *tiled(m))
    if name == "ref512":
        # 4096 cells: 4097 records in every workgroup's copy
        m = ref_model(512, 512, seed=22)
        return (m, *tiled(m))
    if name == "medium512x256":
        m = ref_model(512, 256, seed=23)
        l, t = tiled(m, side=8, nx=50, ny=28)
        t[0] = 0
        return m, l, t
    if name in ("scattered256", "scattered512"):
        # every cell's sites scattered: most targets would change label
        side = int(name[-3:])
        m = ref_model(side, side, seed=24)
        l, t = tiled(m)
        return m, np.random.default_rng(24).permutation(l), t
    if name == "wide1024x32":
        # one row of tiles: the rows above and below a tile are its own
        m = ref_model(1024, 32, seed=25)
        return (m, *tiled(m))
    if name == "tall32x1024":
        # one column of tiles: the columns either side are its own
        m = ref_model(32, 1024, seed=26)
        return (m, *tiled(m))
    raise KeyError(name)


@pytest.mark.parametrize("name,chunks", [("ref256", (70, 3, 1)), ("ref512", (5,)), ("medium512x256", (2, 9)),
                                         ("scattered256", (4,)), ("scattered512", (2,)), ("wide1024x32", (3, 2)),
                                         ("tall32x1024", (3,))])
def test_a_tiled_resident_lattice_matches_across_barriers_launches_and_calls(engine, oracle_lib, name, chunks):
    # plain lattices above the one-workgroup size step with one resident
    # workgroup per tile, passes separated by grid barriers, at most 64
    # steps per launch: runs that cross barriers, launches and calls match
    # the emulation, with tiles of at most one target a lane (priced where
    # drawn) and of more (queued)
    from glazier_amd import HipSimulation

    m, labels, types = coop_case(name)
    steps = sum(chunks)
    print(f"[parity] tiled resident {name}: {m.width}x{m.height}, chunks {chunks}", flush=True)
    gpu = HipSimulation(m, labels, types)
    gpu.set_timing(True)
    for c in chunks:
        gpu.step(c)
    _, launches = gpu.kernel_time("sweep")
    assert launches == sum((c + 63) // 64 for c in chunks), "the tiled resident pass did not run"
    emu = oracle.EmuSimulation(m, labels, types)
    emu.step(steps)
    assert int((gpu.labels() != emu.labels()).sum()) == 0
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert np.array_equal(gpu.volumes(), emu.volumes())
    assert np.array_equal(gpu.surfaces(), emu.surfaces())
    assert gpu.counters() == emu.counters()
    assert gpu.mcs == steps
    gpu.close()
    emu.close()


def test_a_tiled_barrier_that_gives_up_is_reported_and_the_device_stays_usable(engine, oracle_lib):
    # the barrier's wait bounded to no polls from the host: every workgroup of
    # the launch gives up at the first barrier and leaves, so every kernel
    # ends normally; the step reports it, and a new lattice afterwards steps
    # as the emulation does
    from glazier_amd import GhipError, HipSimulation

    m, labels, types = coop_case("ref256")
    print("[parity] tiled resident barrier with no patience", flush=True)
    gpu = HipSimulation(m, labels, types)
    gpu.set_barrier_patience(0)
    with pytest.raises(GhipError) as err:
        gpu.step(2)
    assert "gave up" in str(err.value)
    gpu.close()
    again = HipSimulation(m, labels, types)
    again.step(3)
    emu = oracle.EmuSimulation(m, labels, types)
    emu.step(3)
    assert int((again.labels() != emu.labels()).sum()) == 0
    assert again.counters() == emu.counters()
    again.close()
    emu.close()


def test_span_timing_counts_the_passes_of_each_step_call(engine):
    from glazier_amd import HipSimulation
    from glazier_amd.engine import GhipError

    m, labels, types, _, _, _ = case("surface1024")
    gpu = HipSimulation(m, labels, types)
    gpu.set_timing(2)
    gpu.step(3)
    gpu.step(2)
    gpu.set_timing(False)
    ms, passes = gpu.kernel_time("step")
    assert passes == 4 * 5 and ms > 0.0
    assert gpu.kernel_time("sweep") == (0.0, 0)
    with pytest.raises(GhipError):
        gpu.set_timing(3)
    gpu.close()


@pytest.mark.parametrize("side,temperature,chunks", [(256, 1.6, (1, 2, 5, 12, 20)),
                                                     (256, 1.3, (10, 15)), (1024, 1.6, (2, 3))])
def test_quiet_passes_match_the_emulation(engine, oracle_lib, side, temperature, chunks):
    # a cold plain sheet accepts a copy now and then: many passes accept
    # nothing anywhere, and then the next passes skip the snapshot fold and
    # the previous pass's deltas; the books must stay those of the emulation
    # across every change between quiet and busy passes
    from glazier_amd import HipSimulation
    from glazier_amd.workloads import plain_monolayer

    m, labels, types = plain_monolayer(side, 5, False)
    m.temperature = temperature
    print(f"[parity] quiet {side}^2 at T={temperature}, chunks {chunks}", flush=True)
    gpu = HipSimulation(m, labels, types)
    emu = oracle.EmuSimulation(m, labels, types)
    for c in chunks:
        gpu.step(c)
        emu.step(c)
        assert np.array_equal(gpu.labels(), emu.labels())
        assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
        assert gpu.counters() == emu.counters()
    assert gpu.counters()[1] > 0
    gpu.close()


@pytest.mark.parametrize("name", ["moving", "scattered"])
def test_a_4096_sheet_of_many_rounds_matches_the_emulation(engine, oracle_lib, name):
    # 2048 tiles, more than one round of workgroups: tiles merge their record
    # changes in LDS before the global atomics. The reference's benchmark
    # model accepts 7% of its attempts; with every cell's sites scattered at
    # random nearly every target is queued and a tile's copies touch more
    # labels than its merge table holds
    from glazier_amd import HipSimulation
    from tests.cases import ref_model, tiled

    m = ref_model(4096, 4096, seed=31)
    labels, types = tiled(m)
    if name == "scattered":
        labels = np.random.default_rng(31).permutation(labels)
    steps = 3 if name == "moving" else 1
    print(f"[parity] 4096^2 {name}, {steps} MCS", flush=True)
    gpu = HipSimulation(m, labels, types)
    gpu.step(steps)
    emu = oracle.EmuSimulation(m, labels, types)
    emu.step(steps)
    assert int((gpu.labels() != emu.labels()).sum()) == 0
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.counters() == emu.counters()
    assert gpu.counters()[1] > 0
    gpu.close()
    emu.close()


def test_a_large_lattice_keeps_its_books_across_launches(engine, oracle_lib):
    # the per-step bookkeeping's chain of gated launches at scale: divisions
    # on the first step, deaths every step, and a compaction once a quarter
    # of the labels are dead (step 7)
    from glazier_amd import HipSimulation
    from tests.cases import ref_model, tiled

    m = ref_model(1024, 1024, seed=41, division_volume=64.0, death_rate=0.05)
    labels, types = tiled(m)
    print("[parity] 1024^2 division and death, 8 MCS", flush=True)
    gpu = HipSimulation(m, labels, types)
    emu = oracle.EmuSimulation(m, labels, types)
    for chunk in (1, 5, 2):
        gpu.step(chunk)
        emu.step(chunk)
        assert gpu.n_labels() == emu.n_labels()
        assert np.array_equal(gpu.labels(), emu.labels())
        assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:])
    assert gpu.n_labels() < 26000  # compacted
    gpu.close()

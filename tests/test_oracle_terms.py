"""The terms of the model as the reference's integration tests exercise
them, on the serial oracle: glazier-tests/tests/motility.rs, fields.rs,
volume.rs and volume_terms.rs, one for one (the device halves of motility.rs
and volume_terms.rs, and device versions of the field and volume checks, are
in test_gpu_terms.py). Each reads what a term is for: a remembering cell
travels, a drift carries a cell its way, a secreting sheet fills its field,
a volume runs every term through its third axis.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.blueprint import Blueprint
from glazier_amd.model import CellType, Model, Species
from tests.cases import lone_cell, tiled


# ---------------------------------------------------------------- motility.rs

def motile_model(max_activity, lambda_activity, external, connected):
    # motility.rs:10-33
    c = CellType(target_volume=64.0, lambda_volume=2.0, connected=connected,
                 max_activity=max_activity, lambda_activity=lambda_activity, external=external)
    return Model(width=96, height=96, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                 temperature=10.0, neighbour_order=2, seed=29)


def centroid(labels, m, label=1):
    """Plain mean of the label's site coordinates (motility.rs:47-70)."""
    idx = np.nonzero(np.asarray(labels).reshape(-1) == label)[0]
    w, h = m.width, m.height
    return np.array([(idx % w).mean(), ((idx // w) % h).mean(), (idx // (w * h)).mean()])


def serial_lone(m, lo, hi):
    sim = oracle.SerialSimulation(m, *lone_cell(m, lo, hi))
    return sim


def displacement(m, steps):
    # motility.rs:72-81
    sim = serial_lone(m, 44, 52)
    c0 = centroid(sim.labels(), m)
    sim.step(steps)
    return float(np.linalg.norm(centroid(sim.labels(), m) - c0))


def test_memory_makes_a_cell_travel_further_than_it_jiggles(oracle_lib):
    # motility.rs:83-92
    still = displacement(motile_model(0.0, 0.0, (0.0, 0.0, 0.0), False), 400)
    motile = displacement(motile_model(20.0, 60.0, (0.0, 0.0, 0.0), False), 400)
    assert motile > 3.0 * still, (motile, still)


def test_a_motile_cell_holds_its_volume(oracle_lib):
    # motility.rs:94-105
    sim = serial_lone(motile_model(20.0, 60.0, (0.0, 0.0, 0.0), False), 44, 52)
    sim.step(400)
    assert abs(float(sim.volumes()[1]) - 64.0) < 16.0


@pytest.mark.parametrize("push", [3.0, -3.0])
def test_a_drift_takes_a_cell_the_way_it_points(oracle_lib, push):
    # motility.rs:107-134 (and its reversal)
    m = motile_model(0.0, 0.0, (push, 0.0, 0.0), True)
    sim = serial_lone(m, 44, 52)
    c0 = centroid(sim.labels(), m)
    sim.step(300)
    d = centroid(sim.labels(), m) - c0
    assert np.sign(push) * d[0] > 5.0, d
    if push > 0:
        assert abs(d[1]) < 4.0, d


def test_memory_decays_to_nothing_when_a_cell_stops_being_taken(oracle_lib):
    # motility.rs:136-150
    sim = serial_lone(motile_model(5.0, 60.0, (0.0, 0.0, 0.0), False), 44, 52)
    sim.step(20)
    assert (sim.activity() > 0.0).any(), "a moving cell should remember"
    for _ in range(10):
        sim.decay_activity()
    assert (sim.activity() <= 0.0).all(), "memory should have run out"


def test_a_run_with_neither_term_leaves_the_memory_empty(oracle_lib):
    # motility.rs:152-160
    sim = serial_lone(motile_model(0.0, 0.0, (0.0, 0.0, 0.0), False), 44, 52)
    sim.step(50)
    assert (sim.activity() <= 0.0).all()


# ---------------------------------------------------------------- fields.rs

def field_model(secretion, uptake, decay):
    # fields.rs:8-51
    return Model(width=48, height=48, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), CellType(target_volume=64.0, lambda_volume=2.0)],
                 temperature=10.0, neighbour_order=2, seed=1,
                 species=[Species("IFNg", 0.2, decay, 0.0)],
                 secretion=[[0.0], [secretion]], uptake=[[0.0], [uptake]], chemotaxis=[[0.0], [0.0]])


def test_a_secreting_sheet_fills_the_field(oracle_lib):
    # fields.rs:53-71
    m = field_model(0.1, 0.0, 0.0)
    sim = oracle.SerialSimulation(m, *tiled(m, side=8, nx=3, ny=3))
    assert sim.field(0).sum() == 0.0
    sim.step(20)
    f = sim.field(0)
    assert f.sum() > 100.0
    assert f[sim.labels() == 0].sum() > 0.0, "nothing reached the medium"


def test_decay_bounds_a_secreting_sheet(oracle_lib):
    # fields.rs:73-87
    steady_m, growing_m = field_model(0.1, 0.0, 0.05), field_model(0.1, 0.0, 0.0)
    steady = oracle.SerialSimulation(steady_m, *tiled(steady_m, side=8, nx=3, ny=3))
    growing = oracle.SerialSimulation(growing_m, *tiled(growing_m, side=8, nx=3, ny=3))
    steady.step(80)
    growing.step(80)
    assert steady.field(0).sum() < growing.field(0).sum()


def test_uptake_removes_what_secretion_adds(oracle_lib):
    # fields.rs:89-106
    m = field_model(0.1, 1.0, 0.0)
    sim = oracle.SerialSimulation(m, *tiled(m, side=8, nx=3, ny=3))
    sim.step(30)
    assert (sim.field(0)[sim.labels() != 0] < 1e-9).all()


FIELD_TEXT = """{
    "name": "one field",
    "width": 32, "height": 32,
    "types": [{ "name": "Cell", "target_volume": 64.0, "lambda_volume": 2.0,
                "secretion": {"%s": 0.5}, "uptake": {"IFNg": 0.01} }],
    "fields": [{ "name": "IFNg", "diffusion": 0.2, "decay": 0.001 }],
    "contact": [0.0, 16.0, 16.0, 8.0],
    "temperature": 10.0, "neighbour_order": 2, "seed": 1,
    "steps": 10, "dump_every": 0,
    "initial": { "side": 8, "nx": 2, "ny": 2 },
    "units": { "micron_per_site": 1.0, "minute_per_step": 1.0 }
}"""


def test_a_description_states_its_fields():
    # fields.rs:108-129, through both readers
    from glazier_amd import native

    m = Blueprint.from_json(FIELD_TEXT % "IFNg").model()
    assert [s.name for s in m.species] == ["IFNg"]
    assert m.secretion[0] == [0.0] and m.secretion[1] == [0.5] and m.uptake[1] == [0.01]
    parsed = native.parse(FIELD_TEXT % "IFNg")
    assert parsed["species"] == ["IFNg"]
    assert parsed["secretion"] == [[0.0], [0.5]] and parsed["uptake"][1] == [0.01]


def test_secreting_a_field_the_description_never_declared_is_an_error():
    # fields.rs:131-146, through both readers
    from glazier_amd import native

    with pytest.raises(ValueError, match="IFNgamma"):
        Blueprint.from_json(FIELD_TEXT % "IFNgamma")
    with pytest.raises(ValueError, match="IFNgamma"):
        native.parse(FIELD_TEXT % "IFNgamma")


# ---------------------------------------------------------------- volume.rs

def cube_model(order):
    # volume.rs:8-28
    return Model(width=32, height=32, depth=32, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), CellType(target_volume=512.0, lambda_volume=2.0)],
                 temperature=10.0, neighbour_order=order, seed=13)


def test_a_volume_tiles_into_cubes(oracle_lib):
    # volume.rs:30-38
    m = cube_model(2)
    sim = oracle.SerialSimulation(m, *tiled(m))
    v = sim.volumes()
    assert v.size - 1 == 64 and (v[1:] == 512).all()


def test_the_incremental_delta_matches_a_full_recomputation_in_a_volume(oracle_lib):
    # volume.rs:40-70
    m = cube_model(2)
    sim = oracle.SerialSimulation(m, *tiled(m))
    rng = oracle.Xoshiro(17)
    checked = 0
    for _ in range(5):
        sim.step()
        for _ in range(15):
            lab = sim.labels()
            target, source = rng.below(lab.size), rng.below(lab.size)
            new, old = int(lab[source]), int(lab[target])
            if old == new or (old != 0 and sim.volumes()[old] <= 1):
                continue
            predicted = sim.delta_energy(target, new)
            before = sim.energy()
            after = sim.clone()
            after.apply_copy(target, new)
            measured = after.energy() - before
            after.close()
            assert abs(predicted - measured) < 1e-9, (target, predicted, measured)
            checked += 1
    assert checked > 20


def test_the_bookkeeping_survives_a_run_in_a_volume(oracle_lib):
    # volume.rs:72-80
    m = cube_model(2)
    sim = oracle.SerialSimulation(m, *tiled(m))
    sim.step(20)
    assert np.array_equal(sim.volumes(), sim.recounted_volumes())
    assert np.array_equal(sim.surfaces(), sim.recounted_surfaces())


def deepest(labels, m):
    idx = np.nonzero(np.asarray(labels) != 0)[0]
    return int((idx // (m.width * m.height)).max())


def test_a_cell_reaches_across_the_third_axis(oracle_lib):
    # volume.rs:82-104: a slab eight layers deep spreads through the depth
    m = cube_model(2)
    sim = oracle.SerialSimulation(m, *tiled(m, side=8, nx=4, ny=4, nz=1))
    assert deepest(sim.labels(), m) == 7
    sim.step(30)
    assert deepest(sim.labels(), m) > 7


def test_a_field_diffuses_through_the_depth(oracle_lib):
    # volume.rs:106-131
    m = cube_model(2)
    m.species = [Species("signal", 0.3, 0.0, 0.0)]
    m.secretion, m.uptake, m.chemotaxis = [[0.0], [0.0]], [[0.0], [0.0]], [[0.0], [0.0]]
    sim = oracle.SerialSimulation(m, *tiled(m))
    f = sim.field(0)
    centre = (16 * 32 + 16) * 32 + 16
    f[centre] = 1000.0
    sim.set_field(0, f)
    before = sim.field(0).sum()
    sim.step(20)
    assert abs(sim.field(0).sum() - before) < 1e-6, "diffusion lost mass"
    assert sim.field(0)[(19 * 32 + 16) * 32 + 16] > 0.0, "nothing reached along z"


def test_a_higher_neighbour_order_gives_a_cell_more_boundary(oracle_lib):
    # volume.rs:133-145
    faces_m, corners_m = cube_model(1), cube_model(3)
    faces = oracle.SerialSimulation(faces_m, *tiled(faces_m))
    corners = oracle.SerialSimulation(corners_m, *tiled(corners_m))
    assert corners.surfaces()[1:].mean() > faces.surfaces()[1:].mean()


# ---------------------------------------------------------------- volume_terms.rs

def base(seed):
    # volume_terms.rs:13-33
    return Model(width=32, height=32, depth=32, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), CellType(target_volume=512.0, lambda_volume=2.0)],
                 temperature=10.0, neighbour_order=2, seed=seed)


def test_a_drift_along_the_third_axis_takes_a_cell_that_way(oracle_lib):
    # volume_terms.rs:68-98
    m = base(51)
    m.types[1].external = (0.0, 0.0, 10.0)
    m.types[1].connected = True
    sim = serial_lone(m, 12, 20)
    c0 = centroid(sim.labels(), m)
    sim.step(400)
    d = centroid(sim.labels(), m) - c0
    assert d[2] > 1.5, d
    assert abs(d[0]) < 2.0 and abs(d[1]) < 2.0, d


def test_memory_moves_a_cell_further_in_a_volume_too(oracle_lib):
    # volume_terms.rs:100-132
    def travel(max_activity, lambda_activity):
        m = base(53)
        m.types[1].max_activity = max_activity
        m.types[1].lambda_activity = lambda_activity
        sim = serial_lone(m, 12, 20)
        c0 = centroid(sim.labels(), m)
        sim.step(400)
        return float(np.linalg.norm(centroid(sim.labels(), m) - c0))

    still = travel(0.0, 0.0)
    short_memory = travel(20.0, 60.0)
    motile = travel(100.0, 100.0)
    assert abs(short_memory - still) < 0.1, (short_memory, still)
    assert motile > 3.0 * still, (motile, still)


def test_a_cell_divides_across_its_own_axis_in_a_volume(oracle_lib):
    # volume_terms.rs:134-153
    m = base(57)
    m.types[1].division_volume = 400.0
    m.types[1].target_volume = 900.0
    sim = serial_lone(m, 12, 20)
    assert int((sim.volumes()[1:] > 0).sum()) == 1
    divided, died = sim.divide_and_die()
    assert (divided, died) == (1, 0)
    v = sim.volumes()
    assert int((v[1:] > 0).sum()) == 2
    assert np.array_equal(v, sim.recounted_volumes())
    assert int(v[1]) + int(v[2]) == 512


def test_death_empties_a_volume_at_the_stated_rate(oracle_lib):
    # volume_terms.rs:155-165
    m = base(59)
    m.types[1].death_rate = 1.0
    sim = oracle.SerialSimulation(m, *tiled(m))
    assert int((sim.volumes()[1:] > 0).sum()) == 64
    sim.divide_and_die()
    assert int((sim.volumes()[1:] > 0).sum()) == 0
    assert (sim.labels() == 0).all()


def test_a_length_target_stretches_a_cell_through_the_depth(oracle_lib):
    # volume_terms.rs:167-186
    m = base(61)
    m.types[1].target_length = 30.0
    m.types[1].lambda_length = 1.0
    long = serial_lone(m, 12, 20)
    round_ = serial_lone(base(61), 12, 20)
    long.step(150)
    round_.step(150)
    assert long.recounted_lengths()[1] > round_.recounted_lengths()[1]


def test_a_secreting_cell_fills_a_volume_around_itself(oracle_lib):
    # volume_terms.rs:188-218
    m = base(63)
    m.species = [Species("signal", 0.3, 0.0, 0.0)]
    m.secretion, m.uptake, m.chemotaxis = [[0.0], [0.5]], [[0.0], [0.0]], [[0.0], [0.0]]
    sim = serial_lone(m, 12, 20)
    sim.step(40)
    f = sim.field(0)
    assert f.sum() > 100.0
    assert f[(26 * 32 + 16) * 32 + 16] > 0.0, "nothing reached ten sites away along z"


# ---------------------------------------------------------------- connectivity.rs

def conn_model(connected, temperature, d3=False):
    # connectivity.rs:8-27 (and the volume of 99-117)
    c = CellType(target_volume=512.0 if d3 else 64.0, lambda_volume=1.0, connected=connected)
    side = 32 if d3 else 64
    return Model(width=side, height=side, depth=side if d3 else 1, contact=[0.0, 4.0, 4.0, 4.0],
                 types=[CellType(), c], temperature=temperature, neighbour_order=2, seed=23)


def pieces(labels, m):
    """Pieces the labels fall into under the ring rule (8 neighbours in a
    plane, 26 in a volume; connectivity.rs:30-53), periodic: each site takes
    the smallest site index reachable through same-label ring neighbours."""
    shape = m.shape()
    lab = np.asarray(labels).reshape(shape)
    comp = np.arange(lab.size).reshape(shape)
    offsets = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
               if (dz, dy, dx) != (0, 0, 0) and (len(shape) == 3 or dz == 0)]
    while True:
        before = comp
        for dz, dy, dx in offsets:
            shift = (dy, dx) if len(shape) == 2 else (dz, dy, dx)
            axes = tuple(range(len(shape)))
            nl = np.roll(lab, shift, axis=axes)
            nc = np.roll(comp, shift, axis=axes)
            comp = np.where((nl == lab) & (lab != 0), np.minimum(comp, nc), comp)
        if np.array_equal(comp, before):
            break
    return int(np.unique(comp[lab != 0]).size)


def test_a_tiled_sheet_starts_in_one_piece_per_cell(oracle_lib):
    # connectivity.rs:55-59
    m = conn_model(True, 10.0)
    labels, _ = tiled(m)
    assert pieces(labels, m) == int(labels.max())


def test_the_constraint_keeps_cells_whole_where_nothing_else_does(oracle_lib):
    # connectivity.rs:61-85
    loose_m, whole_m = conn_model(False, 40.0), conn_model(True, 40.0)
    loose = oracle.SerialSimulation(loose_m, *tiled(loose_m))
    whole = oracle.SerialSimulation(whole_m, *tiled(whole_m))
    loose.step(200)
    whole.step(200)
    cells = loose.n_labels() - 1
    assert pieces(loose.labels(), loose_m) > cells, "the unconstrained sheet stayed whole"
    assert pieces(whole.labels(), whole_m) == cells


def test_the_constraint_leaves_the_population_and_the_volume_alone(oracle_lib):
    # connectivity.rs:87-97
    m = conn_model(True, 20.0)
    sim = oracle.SerialSimulation(m, *tiled(m))
    started = sim.n_labels()
    sim.step(100)
    assert sim.n_labels() == started
    assert np.array_equal(sim.volumes(), sim.recounted_volumes())
    assert np.array_equal(sim.surfaces(), sim.recounted_surfaces())


def test_a_volume_keeps_its_cells_whole_too(oracle_lib):
    # connectivity.rs:99-117
    m = conn_model(True, 40.0, d3=True)
    sim = oracle.SerialSimulation(m, *tiled(m))
    sim.step(40)
    assert pieces(sim.labels(), m) == sim.n_labels() - 1


# ---------------------------------------------------------------- chemotaxis.rs

def chem_model(sensitivity):
    # chemotaxis.rs:7-35
    return Model(width=64, height=64, contact=[0.0, 16.0, 16.0, 8.0],
                 types=[CellType(), CellType(target_volume=64.0, lambda_volume=2.0)],
                 temperature=10.0, neighbour_order=2, seed=4,
                 species=[Species("attractant", 0.0, 0.0, 0.0)],
                 secretion=[[0.0], [0.0]], uptake=[[0.0], [0.0]], chemotaxis=[[0.0], [sensitivity]])


def chem_drift(sensitivity, steps=300):
    # chemotaxis.rs:38-83: the ramp slope * x is imposed before every step
    m = chem_model(sensitivity)
    sim = oracle.SerialSimulation(m, *lone_cell(m, 24, 32))
    ramp = np.tile(np.arange(64, dtype=np.float64), 64)
    x0 = centroid(sim.labels(), m)[0]
    for _ in range(steps):
        sim.set_field(0, ramp)
        sim.step()
    return centroid(sim.labels(), m)[0] - x0


@pytest.mark.parametrize("sensitivity", [4.0, -4.0, 0.0])
def test_a_cell_follows_the_gradient_its_sensitivity_states(oracle_lib, sensitivity):
    # chemotaxis.rs:85-107: up the ramp, down it, or nowhere
    moved = chem_drift(sensitivity)
    if sensitivity > 0:
        assert moved > 3.0, moved
    elif sensitivity < 0:
        assert moved < -3.0, moved
    else:
        assert abs(moved) < 3.0, moved


def test_a_description_states_its_sensitivities():
    # chemotaxis.rs:121-139, through both readers
    from glazier_amd import native

    text = FIELD_TEXT.replace('"secretion": {"%s": 0.5}, "uptake": {"IFNg": 0.01}',
                              '"chemotaxis": {"IFNg": 300.0}')
    m = Blueprint.from_json(text).model()
    assert m.chemotaxis[0] == [0.0] and m.chemotaxis[1] == [300.0]
    assert native.parse(text)["chemotaxis"] == [[0.0], [300.0]]


# ---------------------------------------------------------------- length.rs

def length_model(target_length, lambda_length):
    # length.rs:7-27
    c = CellType(target_volume=64.0, lambda_volume=2.0, target_length=target_length,
                 lambda_length=lambda_length)
    return Model(width=48, height=48, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
                 temperature=10.0, neighbour_order=2, seed=3)


def test_moment_bookkeeping_survives_a_run(oracle_lib):
    # length.rs:73-88
    m = length_model(20.0, 1.0)
    sim = oracle.SerialSimulation(m, *tiled(m))
    sim.step(80)
    assert np.allclose(sim.lengths()[1:], sim.recounted_lengths()[1:], rtol=0.0, atol=1e-6)


def test_a_length_target_stretches_the_cells(oracle_lib):
    # length.rs:90-109
    round_m, long_m = length_model(0.0, 0.0), length_model(24.0, 1.0)
    round_ = oracle.SerialSimulation(round_m, *tiled(round_m))
    long = oracle.SerialSimulation(long_m, *tiled(long_m))
    round_.step(300)
    long.step(300)
    assert long.recounted_lengths()[1:].mean() > round_.recounted_lengths()[1:].mean()


# ---------------------------------------------------------------- adhesion.rs

def adhesion_text(binding, second_presents, drop_block=False, molecule="cadherin", shape="[%s]"):
    # adhesion.rs:6-29
    block = '' if drop_block else '"adhesion": { "molecules": ["cadherin"], "binding": %s },' % (
        shape % repr(float(binding)))
    return """{
        "name": "two types with one molecule",
        "width": 64, "height": 64,
        "types": [
            { "name": "Sticky", "target_volume": 64.0, "lambda_volume": 2.0,
              "presents": { "%s": 1.0 } },
            { "name": "Other", "target_volume": 64.0, "lambda_volume": 2.0,
              "presents": { "cadherin": %r } }
        ],
        %s
        "contact": [0.0, 16.0, 16.0, 16.0, 10.0, 10.0, 16.0, 10.0, 10.0],
        "temperature": 10.0, "neighbour_order": 2, "seed": 1,
        "steps": 10, "dump_every": 0,
        "initial": { "side": 8, "nx": 4, "ny": 4 },
        "units": { "micron_per_site": 1.0, "minute_per_step": 1.0 }
    }""" % (molecule, float(second_presents), block)


def written_out_text(cell_cell):
    # adhesion.rs:61-82
    return adhesion_text(0.0, 0.0).replace(
        '"contact": [0.0, 16.0, 16.0, 16.0, 10.0, 10.0, 16.0, 10.0, 10.0]',
        '"contact": [0.0, 16.0, 16.0, 16.0, %r, %r, 16.0, %r, %r]' % ((cell_cell,) * 4))


def contact_of(text, a, b):
    from glazier_amd import native

    bp = Blueprint.from_json(text)
    n = len(bp.type_names)
    c = bp.effective_contact()
    assert list(native.parse(text)["contact"]) == pytest.approx(c)  # both readers fold alike
    return c[a * n + b]


def test_binding_comes_off_the_contact_energy_between_the_types_that_present_it():
    # adhesion.rs:36-45
    text = adhesion_text(4.0, 1.0)
    assert abs(contact_of(text, 1, 1) - 6.0) < 1e-12
    assert abs(contact_of(text, 1, 2) - 6.0) < 1e-12
    assert abs(contact_of(text, 0, 1) - 16.0) < 1e-12


def test_a_type_that_presents_nothing_keeps_its_stated_energies():
    # adhesion.rs:47-53
    text = adhesion_text(4.0, 0.0)
    assert abs(contact_of(text, 1, 1) - 6.0) < 1e-12
    assert abs(contact_of(text, 1, 2) - 10.0) < 1e-12
    assert abs(contact_of(text, 2, 2) - 10.0) < 1e-12


def test_no_binding_leaves_the_matrix_as_the_description_states_it():
    # adhesion.rs:55-59
    bp = Blueprint.from_json(adhesion_text(0.0, 1.0))
    assert bp.effective_contact() == bp.contact


def test_a_molecule_description_runs_as_the_matrix_it_folds_to(oracle_lib):
    # adhesion.rs:84-103: exact, the same lattice either way
    folded = Blueprint.from_json(adhesion_text(4.0, 1.0))
    direct = Blueprint.from_json(written_out_text(6.0))
    assert folded.effective_contact() == direct.effective_contact()
    runs = []
    for bp in (folded, direct):
        m, labels, types = bp.simulation()
        sim = oracle.SerialSimulation(m, labels, types)
        sim.step(150)
        runs.append((sim.labels(), sim.volumes()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("text,message", [
    (adhesion_text(4.0, 1.0, molecule="integrin"), "integrin"),
    (adhesion_text(4.0, 1.0, drop_block=True), "declares none"),
    (adhesion_text(4.0, 1.0, shape="[%s, 1.0]"), "binding matrix"),
])
def test_bad_adhesion_descriptions_are_refused_by_both_readers(text, message):
    # adhesion.rs:112-136
    from glazier_amd import native

    with pytest.raises(ValueError, match=message):
        Blueprint.from_json(text)
    with pytest.raises(ValueError, match=message):
        native.parse(text)

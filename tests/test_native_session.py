"""The session surface (glazier_amd.native) against the reference's own
checks of glazier._native (tests/test_native_engine.py:27-48) and the
GpuSession members (glazier-py/src/lib.rs:184-241).

parse() is the engine's reader, so on every shipped blueprint it must agree
with the Python reader as the binary's --check does; that part needs no
device. The runs need the device and are marked gpu.
"""

from pathlib import Path

import numpy as np
import pytest

from glazier_amd.blueprint import Blueprint
from tests.test_blueprint_conformance import BLUEPRINTS, assert_agrees

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden" / "blueprints"
MODEL = GOLDEN / "monolayer.json"


@pytest.mark.parametrize("path", BLUEPRINTS, ids=lambda p: p.stem)
def test_parse_agrees_with_the_python_reader(path):
    from glazier_amd import native

    parsed = native.parse(path.read_text())
    bp = Blueprint.load(path)
    assert_agrees(bp, parsed)
    assert parsed["depth"] == bp.depth


def test_parse_refuses_a_bad_description():
    from glazier_amd import native

    with pytest.raises(ValueError, match="width"):
        native.parse('{"name": "x"}')


@pytest.mark.gpu
def test_a_short_run_returns_a_lattice_of_the_stated_shape(engine):
    # test_native_engine.py:27-37
    from glazier_amd import native

    text = MODEL.read_text()
    labels, width, height, depth = native.run(text, 5)
    field = np.asarray(labels, dtype=np.uint32).reshape(depth, height, width)
    parsed = native.parse(text)
    assert field.shape == (parsed["depth"], parsed["height"], parsed["width"])
    side, nx, ny, nz = parsed["initial"]
    assert np.unique(field[field > 0]).size == nx * ny * nz


@pytest.mark.gpu
def test_volumes_come_back_per_cell(engine):
    # test_native_engine.py:40-48
    from glazier_amd import native

    text = MODEL.read_text()
    volumes = native.run_volumes(text, 5)
    parsed = native.parse(text)
    side, nx, ny, nz = parsed["initial"]
    assert len(volumes) == nx * ny * nz + 1
    sites = parsed["width"] * parsed["height"] * parsed["depth"]
    assert sum(volumes[1:]) == sites


@pytest.mark.gpu
def test_a_session_steps_reads_back_and_takes_a_new_field(engine):
    # the GpuSession members, on the nematic blueprint (its field beside it)
    from glazier_amd import native

    text = (GOLDEN / "monolayer-nematic.json").read_text()
    s = native.GpuSession(text, base=str(GOLDEN))
    assert s.mcs == 0
    s.step(3)
    s.step(2)
    assert s.mcs == 5
    labels = np.asarray(s.labels(), dtype=np.uint32)
    assert labels.size == 128 * 128 and np.unique(labels[labels > 0]).size == 256
    flat = np.tile([1.0, 0.0], 128 * 128).tolist()
    s.set_nematic_field(flat)
    s.step(1)
    assert s.mcs == 6
    with pytest.raises(ValueError, match="expected 32768 numbers"):
        s.set_nematic_field(flat[:-2])
    s.close()


# tests/test_volume_measurement.py of the reference: a volume is measured on
# its middle plane

@pytest.fixture(scope="module")
def volume_field():
    from glazier_amd import native

    text = (GOLDEN / "volume.json").read_text()
    labels, width, height, depth = native.run(text, 5)
    return np.asarray(labels, dtype=np.uint32).reshape(depth, height, width)


@pytest.mark.gpu
def test_a_volume_comes_back_three_dimensional(engine, volume_field):
    bp = Blueprint.load(GOLDEN / "volume.json")
    assert bp.dimensions == 3
    assert volume_field.shape == (bp.depth, bp.height, bp.width)


@pytest.mark.gpu
def test_a_slice_is_one_layer_of_it(engine, volume_field):
    from glazier_amd import mid_plane

    plane = mid_plane(volume_field)
    assert plane.shape == volume_field.shape[1:]
    assert np.array_equal(plane, volume_field[volume_field.shape[0] // 2])


def test_a_plane_is_returned_unchanged():
    from glazier_amd import mid_plane

    flat = np.arange(12, dtype=np.uint32).reshape(3, 4)
    assert np.array_equal(mid_plane(flat), flat)


@pytest.mark.gpu
def test_the_measurements_run_on_a_sliced_volume(engine, volume_field):
    from glazier_amd import director, mid_plane

    plane = mid_plane(volume_field)
    shapes = director.cell_shapes(plane)
    assert shapes, "the slice held no cells"
    q = director.coarse_grain(director.q_field(plane, shapes), 4.0)
    charge = sum(c for _, _, c in director.defect_charges(director.director_angle(q)))
    # a slice of a periodic volume is periodic in both axes, so its charges cancel
    assert abs(charge) < 1e-9

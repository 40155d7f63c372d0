"""The session surface (glazier_amd.native) held to its contract in the
reference's Python module (glazier-py/src/lib.rs:147-241):

* run(text, steps) -> (labels flat with x fastest, width, height, depth);
* run_volumes(text, steps) -> one volume per label, index 0 the medium;
* GpuSession(text, base): step / labels / set_nematic_field / mcs, with a
  ValueError naming the expected count for a field of the wrong length.

Both runs step the HIP engine from the same description and seed, so the
volumes of one must be the label counts of the other's lattice. The device
runs are marked gpu; parse() needs no device.
"""

from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden" / "blueprints"

# (blueprint, steps, sheet fills the lattice)
RUNS = [
    ("monolayer.json", 5, True),
    ("monolayer-surface.json", 4, True),
    ("volume.json", 3, True),
    ("infection-volume.json", 3, False),
]


def test_parse_refuses_a_bad_description():
    from glazier_amd import native

    with pytest.raises(ValueError, match="width"):
        native.parse('{"name": "x"}')


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps,confluent", RUNS, ids=[r[0] for r in RUNS])
def test_run_and_run_volumes_agree_on_one_lattice(engine, name, steps, confluent):
    from glazier_amd import native

    text = (GOLDEN / name).read_text()
    parsed = native.parse(text)
    side, nx, ny, nz = parsed["initial"]
    flat, width, height, depth = native.run(text, steps)
    assert (width, height, depth) == (parsed["width"], parsed["height"], parsed["depth"])
    labels = np.asarray(flat, dtype=np.uint32)
    assert labels.size == width * height * depth
    volumes = np.asarray(native.run_volumes(text, steps), dtype=np.int64)
    cells = nx * ny * nz
    assert volumes.size == cells + 1
    # the same seed and steps give the same lattice, so the volumes are its counts
    assert np.array_equal(np.bincount(labels, minlength=cells + 1), volumes)
    assert volumes.sum() == labels.size
    if confluent:
        assert volumes[0] == 0 and np.count_nonzero(volumes[1:]) == cells


@pytest.mark.gpu
def test_a_session_keeps_its_run_and_takes_a_new_field(engine):
    from glazier_amd import native

    text = (GOLDEN / "monolayer-nematic.json").read_text()
    s = native.GpuSession(text, base=str(GOLDEN))
    try:
        assert s.mcs == 0
        for chunk in (3, 2):
            s.step(chunk)
        assert s.mcs == 5
        labels = np.asarray(s.labels(), dtype=np.uint32)
        assert labels.size == 128 * 128 and np.unique(labels[labels > 0]).size == 256
        along_x = [1.0, 0.0] * (128 * 128)
        s.set_nematic_field(along_x)
        s.step(1)
        assert s.mcs == 6
        with pytest.raises(ValueError, match="expected 32768 numbers, got 32766"):
            s.set_nematic_field(along_x[:-2])
    finally:
        s.close()


@pytest.mark.gpu
def test_a_volume_run_is_measured_on_its_middle_plane(engine):
    from glazier_amd import director, native

    text = (GOLDEN / "volume.json").read_text()
    flat, width, height, depth = native.run(text, 5)
    field = np.asarray(flat, dtype=np.uint32).reshape(depth, height, width)
    assert depth > 1
    plane = director.mid_plane(field)
    assert np.array_equal(plane, field[depth // 2])
    shapes = director.cell_shapes(field)
    assert shapes and {s.label for s in shapes} <= set(np.unique(plane).tolist())
    theta = director.director_angle(director.coarse_grain(director.q_field(field, shapes), 4.0))
    # the middle plane of a periodic volume is a torus: its winding numbers sum to zero
    assert abs(sum(c for _, _, c in director.defect_charges(theta))) < 1e-9


def test_a_plane_is_its_own_middle_plane():
    from glazier_amd import mid_plane

    flat = np.arange(20, dtype=np.uint32).reshape(4, 5)
    assert mid_plane(flat) is flat or np.array_equal(mid_plane(flat), flat)

"""Tissue measurements (glazier_amd.director) against closed forms: a
rectangle's second moments, the order of aligned and crossed cells, and the
charges of an imposed +1/2 / -1/2 disclination pair."""

import numpy as np

from glazier_amd import director


def rect(h, w, y0, x0, ch, cw, label=1, field=None):
    f = np.zeros((h, w), dtype=np.uint32) if field is None else field
    f[y0:y0 + ch, x0:x0 + cw] = label
    return f


def test_a_rectangle_has_the_closed_form_ellipse():
    # population variance of k consecutive integers is (k^2 - 1) / 12
    shapes = director.cell_shapes(rect(32, 32, 10, 5, 4, 12))
    assert len(shapes) == 1
    s = shapes[0]
    cxx, cyy = (12 ** 2 - 1) / 12, (4 ** 2 - 1) / 12
    assert s.area == 48 and s.theta == 0.0
    assert abs(s.eccentricity - (cxx - cyy) / (cxx + cyy)) < 1e-12
    assert s.centre == (10.5, 11.5)
    tall = director.cell_shapes(rect(32, 32, 5, 10, 12, 4))[0]
    assert abs(tall.theta - np.pi / 2) < 1e-12


def test_a_diagonal_cell_points_at_a_quarter_turn():
    # a one-pixel diagonal: c_xx = c_yy = c_xy, so theta = pi/4 and e = 1
    f = np.zeros((32, 32), dtype=np.uint32)
    for k in range(10):
        f[5 + k, 5 + k] = 1
    s = director.cell_shapes(f)[0]
    assert abs(s.theta - np.pi / 4) < 1e-12 and abs(s.eccentricity - 1.0) < 1e-12


def test_small_cells_are_skipped():
    f = rect(16, 16, 2, 2, 1, 3)  # 3 pixels
    assert director.cell_shapes(f) == []


def test_order_of_aligned_and_crossed_cells():
    f = rect(64, 64, 4, 4, 3, 12, 1)
    rect(64, 64, 20, 30, 3, 12, 2, f)
    order, angle = director.global_order(director.cell_shapes(f))
    assert abs(order - 1.0) < 1e-12 and abs(angle) < 1e-12
    rect(64, 64, 40, 4, 12, 3, 3, f)
    rect(64, 64, 40, 30, 12, 3, 4, f)
    order, _ = director.global_order(director.cell_shapes(f))
    assert order < 1e-12


def test_a_disclination_pair_is_found_with_its_charges():
    n = 96
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    # +1/2 at (40.5, 48.5), -1/2 at (56.5, 48.5): winding pi and -pi around them
    theta = 0.5 * np.arctan2(y - 48.5, x - 40.5) - 0.5 * np.arctan2(y - 48.5, x - 56.5)
    found = director.defect_charges(theta)
    charges = sorted(c for _, _, c in found)
    assert len(found) == 2
    assert abs(charges[0] + 0.5) < 1e-9 and abs(charges[1] - 0.5) < 1e-9
    where = {round(c * 2): (x_, y_) for x_, y_, c in found}
    assert where[1] == (40, 48) and where[-1] == (56, 48)
    assert director.defect_charges(np.full((32, 32), 0.3)) == []


def test_a_tiled_sheet_measures_as_squares():
    from glazier_amd.model import Model, tiled_labels

    m = Model(width=64, height=64)
    labels = tiled_labels(m, 8, 8, 8).reshape(64, 64)
    row = director.measure(labels, 64.0)
    assert row["cells"] == 64 and row["mean_area"] == 64.0 and row["area_rms_from_target"] == 0.0
    assert row["mean_eccentricity"] < 1e-12
    assert row["defects"] == 0 and row["net_charge"] == 0.0


def test_the_serial_oracle_reproduces_the_published_cpu_column():
    # README.md:153-162, glazier CPU column of the engine comparison:
    # monolayer-physicell.json, 200 steps, seed 1, measured with sigma 4.
    # The same stream and the same measurements give the same digits, so
    # this pins the serial oracle's whole trajectory and the measurements.
    from pathlib import Path

    import oracle
    from glazier_amd.blueprint import Blueprint

    path = Path(__file__).resolve().parent / "golden" / "blueprints" / "monolayer-physicell.json"
    bp = Blueprint.load(path)
    model, labels, types = bp.simulation()
    sim = oracle.SerialSimulation(model, labels, types)
    sim.step(bp.steps)
    row = director.measure(sim.labels().reshape(model.shape()), bp.types[0].target_volume, 4.0)
    sim.close()
    assert row["cells"] == 256 and row["mean_area"] == 64.0
    assert f"{row['area_rms_from_target']:.2f}" == "1.58"
    assert f"{row['mean_eccentricity']:.3f}" == "0.397"
    assert f"{row['nematic_order']:.3f}" == "0.064"
    assert row["defects"] == 56 and row["net_charge"] == 0.0


# ---- pinned to the reference's measurements of fixed label fields
# (tests/golden/director_vectors.npz, written by the reference's
# python/glazier/director.py through tests/golden/make_reference_vectors.py)

def _vectors():
    from pathlib import Path

    path = Path(__file__).resolve().parent / "golden" / "director_vectors.npz"
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    names = sorted({k.split("__")[0] for k in data})
    return names, data


def _angle_gap(a, b):
    """Distance of two angles taken mod pi."""
    d = np.mod(np.asarray(a) - np.asarray(b) + 0.5 * np.pi, np.pi) - 0.5 * np.pi
    return np.abs(d)


def test_measurements_match_the_reference_on_fixed_fields():
    names, v = _vectors()
    assert names == ["brick", "ragged", "volume"]
    for name in names:
        labels = v[f"{name}__labels"]
        plane = director.mid_plane(labels)
        assert np.array_equal(plane, v[f"{name}__mid_plane"]), name
        shapes = director.cell_shapes(labels)
        want = v[f"{name}__shapes"]
        got = np.array([[s.label, s.area, s.centre[0], s.centre[1], s.theta, s.eccentricity]
                        for s in shapes])
        assert got.shape == want.shape, name
        assert np.array_equal(got[:, :2], want[:, :2]), name  # labels and areas exactly
        assert np.allclose(got[:, 2:4], want[:, 2:4], rtol=0, atol=1e-9), name
        assert _angle_gap(got[:, 4], want[:, 4]).max() < 1e-9, name
        assert np.allclose(got[:, 5], want[:, 5], rtol=0, atol=1e-9), name
        q = director.q_field(labels, shapes)
        assert np.allclose(q, v[f"{name}__q"], rtol=0, atol=1e-9), name
        qc = director.coarse_grain(q, 4.0)
        assert np.allclose(qc, v[f"{name}__q_coarse"], rtol=0, atol=1e-9), name
        theta = director.director_angle(qc)
        assert _angle_gap(theta, v[f"{name}__theta"]).max() < 1e-6, name
        order = 2.0 * np.hypot(qc[..., 0], qc[..., 1])
        assert np.allclose(order, v[f"{name}__order"], rtol=0, atol=1e-9), name
        glob = director.global_order(shapes)
        assert abs(glob[0] - v[f"{name}__global_order"][0]) < 1e-9, name
        assert _angle_gap(glob[1], v[f"{name}__global_order"][1]) < 1e-9, name
        got_defects = np.array(director.defect_charges(theta), dtype=np.float64).reshape(-1, 3)
        want_defects = v[f"{name}__defects"]
        assert got_defects.shape == want_defects.shape, name
        assert np.array_equal(got_defects[:, :2], want_defects[:, :2]), name
        assert np.allclose(got_defects[:, 2], want_defects[:, 2], rtol=0, atol=1e-9), name

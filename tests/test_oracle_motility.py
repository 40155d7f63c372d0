"""The activity geometric mean of both oracles against the reference's own
known answers (glazier-core/src/motility.rs:81-140) and the closed form.

The serial oracle restates Activity::neighbourhood_mean (motility.rs:56-78)
in f64; the emulation prices the same mean in f32 with the engine's own
log2/exp2 (held to the C library in test_oracle_math.py). Both are asked on
the reference's two-cell lattice: an 8 x 8 plane, Moore neighbours, label 1
on x < 4 and medium elsewhere (motility.rs:85-94).
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model

W = H = 8


def two_cell():
    m = Model(width=W, height=H, contact=[0.0, 16.0, 16.0, 8.0],
              types=[CellType(), CellType(target_volume=32.0, lambda_volume=1.0, max_activity=20.0,
                                          lambda_activity=10.0)],
              temperature=10.0, neighbour_order=2, seed=1)
    labels = np.zeros((H, W), dtype=np.uint32)
    labels[:, :4] = 1
    return m, labels.reshape(-1), np.array([1, 1], dtype=np.uint8)


def site(x, y):
    return y * W + x


def means(activity, at, label=1):
    m, labels, types = two_cell()
    serial = oracle.SerialSimulation(m, labels, types)
    serial.set_activity(activity)
    emu = oracle.EmuSimulation(m, labels, types, activity=activity)
    return serial.activity_mean(at, label), emu.activity_mean(at, label)


def cell_activity(value):
    m, labels, _ = two_cell()
    return np.where(labels == 1, value, 0.0).astype(np.float64)


def test_a_forgotten_neighbour_takes_the_mean_to_nothing():
    # motility.rs:106-113
    act = np.zeros(W * H)
    act[site(1, 4)] = 10.0
    assert means(act, site(1, 4)) == (0.0, 0.0)


def test_a_uniformly_remembered_cell_reads_its_own_value():
    # motility.rs:115-126
    s, e = means(cell_activity(7.0), site(1, 4))
    assert abs(s - 7.0) < 1e-9
    assert abs(e - 7.0) <= 2 * np.spacing(np.float32(7.0))


def test_the_mean_sits_between_the_values_it_reads():
    # motility.rs:128-140, and the closed form: all eight Moore neighbours of
    # (1, 4) hold 4 and the site 16, so the mean is (16 * 4^8)^(1/9) = 2^(20/9)
    act = cell_activity(4.0)
    act[site(1, 4)] = 16.0
    s, e = means(act, site(1, 4))
    assert 4.0 < s < 16.0 and 4.0 < e < 16.0
    exact = 2.0 ** (20.0 / 9.0)
    assert abs(s - exact) < 1e-12
    assert abs(e - exact) <= 3 * np.spacing(np.float32(exact))


def test_only_the_label_s_own_neighbours_count():
    # at (3, 4) the column x = 4 is medium: the mean runs over the site and
    # its five neighbours of label 1; a medium neighbour at zero does not
    # zero it, and asked about the medium's label the site counts alone
    act = cell_activity(2.0)
    act[site(2, 3)] = 32.0
    s, e = means(act, site(3, 4))
    exact = 2.0 ** ((1 + 1 + 1 + 1 + 1 + 5) / 6.0)
    assert abs(s - exact) < 1e-12
    assert abs(e - exact) <= 3 * np.spacing(np.float32(exact))
    s0, e0 = means(act, site(3, 4), label=0)
    assert s0 == 0.0 and e0 == 0.0  # the medium around it remembers nothing


def test_a_site_with_no_memory_reads_nothing_whatever_its_neighbours():
    act = cell_activity(9.0)
    act[site(1, 4)] = 0.0
    assert means(act, site(1, 4)) == (0.0, 0.0)


@pytest.mark.parametrize("value", [1.0, 3.0, 100.0, 1e-3])
def test_a_lone_remembering_site_reads_itself(value):
    # a site whose cell has no other site nearby reads its own value
    m, _, types = two_cell()
    labels = np.zeros(W * H, dtype=np.uint32)
    labels[site(5, 5)] = 1
    act = np.zeros(W * H)
    act[site(5, 5)] = value
    serial = oracle.SerialSimulation(m, labels, types)
    serial.set_activity(act)
    emu = oracle.EmuSimulation(m, labels, types, activity=act)
    assert abs(serial.activity_mean(site(5, 5), 1) - value) < 1e-12 * value
    assert abs(emu.activity_mean(site(5, 5), 1) - value) <= 3 * np.spacing(np.float32(value))

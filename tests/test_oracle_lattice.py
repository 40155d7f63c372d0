"""The reference's lattice unit tests (glazier-core/src/lattice.rs:224-296)
restated on the serial oracle, the emulation and the tiler.

Neighbour counts are read as the surface of a lone one-site cell in medium:
every neighbour of that site holds another label, so its bond count is the
size of the neighbourhood (lattice.rs:255-268: 8, 4, 8 in a plane for orders
2, 1, 3; 6, 18, 26 in a volume). Periodic indexing is read the same way at
the lattice's corner, where every neighbour lies across an edge.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, Model, default_cell_types, tiled_labels

COUNTS = {(1, 2): 4, (2, 2): 8, (3, 2): 8, (1, 3): 6, (2, 3): 18, (3, 3): 26}


def lone_site(w, h, d, order, site):
    m = Model(width=w, height=h, depth=d, contact=[0.0, 16.0, 16.0, 8.0],
              types=[CellType(), CellType(target_volume=1.0, lambda_volume=2.0)],
              temperature=10.0, neighbour_order=order, seed=1)
    labels = np.zeros(w * h * d, dtype=np.uint32)
    labels[site] = 1
    return m, labels, default_cell_types(2)


@pytest.mark.parametrize("order,dims", sorted(COUNTS))
@pytest.mark.parametrize("where", ["middle", "corner"])
def test_a_lone_site_has_the_neighbourhood_for_its_surface(oracle_lib, order, dims, where):
    w, h, d = (8, 6, 1) if dims == 2 else (6, 6, 4)
    if where == "corner":
        site = 0
    else:
        site = ((d // 2) * h + h // 2) * w + w // 2 if dims == 3 else (h // 2) * w + w // 2
    m, labels, types = lone_site(w, h, d, order, site)
    want = COUNTS[(order, dims)]
    for sim in (oracle.SerialSimulation(m, labels, types), oracle.EmuSimulation(m, labels, types)):
        assert int(sim.surfaces()[1]) == want
        assert int(sim.volumes()[1]) == 1
        sim.close()


def test_tiling_lays_down_equal_cells_in_the_plane():
    # lattice.rs:271-283: a 12 x 8 plane tiled with side 4 holds six cells of
    # 16 sites and no medium
    m = Model(width=12, height=8, depth=1, types=[CellType(), CellType()])
    labels = tiled_labels(m, 4, 99, 99)
    counts = np.bincount(labels)
    assert counts[0] == 0
    assert labels.max() == 6 and (counts[1:] == 16).all()


def test_tiling_lays_down_equal_cells_in_a_volume():
    # lattice.rs:285-295: an 8^3 volume tiled with side 4 holds eight cells of 64
    m = Model(width=8, height=8, depth=8, types=[CellType(), CellType()])
    labels = tiled_labels(m, 4, 99, 99, 99)
    counts = np.bincount(labels)
    assert labels.max() == 8 and (counts[1:] == 64).all()


def test_tiling_counts_labels_x_fastest():
    # Lattice::tile_grid (lattice.rs:180-194): label = 1 + bx + nx * by
    m = Model(width=12, height=8, depth=1, types=[CellType(), CellType()])
    labels = tiled_labels(m, 4, 3, 2).reshape(8, 12)
    assert labels[0, 0] == 1 and labels[0, 4] == 2 and labels[0, 8] == 3
    assert labels[4, 0] == 4 and labels[7, 11] == 6


@pytest.mark.gpu
@pytest.mark.parametrize("order,dims", sorted(COUNTS))
def test_the_device_counts_the_same_neighbourhood(engine, order, dims):
    from glazier_amd import HipSimulation

    w, h, d = (8, 6, 1) if dims == 2 else (6, 6, 4)
    for site in (0, w * h * d - 1):
        m, labels, types = lone_site(w, h, d, order, site)
        gpu = HipSimulation(m, labels, types)
        assert int(gpu.surfaces()[1]) == COUNTS[(order, dims)]
        assert int(gpu.volumes()[1]) == 1
        gpu.close()

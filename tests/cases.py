"""Named models and starting lattices shared by the parity tests.

Every case builds (model, labels, cell_types[, fields, activity]) from the
reference's own test models (glazier-tests/tests/*.rs) or the measured
workloads, so the GPU and its oracles start from the same state.
"""

from __future__ import annotations

import numpy as np

from glazier_amd.model import CellType, Model, Species, default_cell_types, tiled_labels
from glazier_amd.workloads import infection_model, infection_sheet, plain_monolayer


def ref_model(w=64, h=64, d=1, seed=1, order=2, contact=(0.0, 16.0, 16.0, 8.0), temperature=10.0,
              **cell):
    """gpu_matches_cpu.rs:14-33: medium-cell 16, cell-cell 8, V_t 64, lambda_V 2."""
    c = CellType(target_volume=64.0, lambda_volume=2.0, **cell)
    return Model(width=w, height=h, depth=d, contact=list(contact), types=[CellType(), c],
                 temperature=temperature, neighbour_order=order, seed=seed)


def tiled(model, side=8, nx=None, ny=None, nz=None):
    nx = model.width // side if nx is None else nx
    ny = model.height // side if ny is None else ny
    nz = (model.depth // side if model.depth > 1 else 1) if nz is None else nz
    labels = tiled_labels(model, side, nx, ny, nz)
    return labels, default_cell_types(int(labels.max()) + 1)


def lone_cell(model, lo, hi):
    """One cell of type 1 on a medium lattice, over [lo, hi) on every axis."""
    shape = model.shape()
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    inside = np.ones(shape, dtype=bool)
    for g in grids:
        inside &= (g >= lo) & (g < hi)
    labels = inside.astype(np.uint32).reshape(-1)
    return labels, default_cell_types(2)


def case(name):
    """(model, labels, types, fields, activity, steps) for a named case."""
    fields = activity = None
    if name == "plain128":
        m, l, t = plain_monolayer(128, 1, False)
        return m, l, t, None, None, 20
    if name == "surface256":
        m, l, t = plain_monolayer(256, 2, True)
        m.types[1].target_surface = 90.0  # near an 8x8 square's 92 bonds, so copies land
        return m, l, t, None, None, 10
    if name == "surface1024":
        m, l, t = plain_monolayer(1024, 3, True)
        return m, l, t, None, None, 3
    if name == "scattered1024":
        # the plain fast path on 1024^2 with every cell's 64 sites scattered
        # at random: nearly every target is queued (more than one per lane),
        # and a tile's accepted copies touch more labels than its merge table
        # holds
        m, l, t = plain_monolayer(1024, 4, False)
        l = np.random.default_rng(4).permutation(l)
        return m, l, t, None, None, 3
    if name == "fast_medium_flat":
        # plain fast path with medium gaps; the medium prices as type 1
        m = ref_model(128, 128, seed=12)
        l, t = tiled(m, side=8, nx=12, ny=12)
        return m, l, t, None, None, 15
    if name == "fast_medium_typed":
        # plain fast path where the medium prices as the medium (cell_type[0] = 0)
        m = ref_model(128, 128, seed=13)
        l, t = tiled(m, side=8, nx=12, ny=12)
        t[0] = 0
        return m, l, t, None, None, 15
    if name == "fast_lone_typed":
        m = ref_model(256, 128, seed=14)
        l, t = lone_cell(m, 40, 48)
        t[0] = 0
        return m, l, t, None, None, 20
    if name == "ref64":
        m = ref_model(seed=5)
        return (m, *tiled(m), None, None, 30)
    if name == "medium_gaps":
        m = ref_model(12, 20, seed=6)
        l, t = tiled(m, side=4, nx=2, ny=3)
        return m, l, t, None, None, 25
    if name == "rect64x32":
        m = ref_model(64, 32, seed=8)
        return (m, *tiled(m), None, None, 20)
    if name == "von_neumann":
        m = ref_model(64, 64, seed=9, order=1)
        return (m, *tiled(m), None, None, 20)
    if name == "float_energies":
        m = ref_model(64, 64, seed=10, contact=(0.0, 16.5, 16.5, 7.25), temperature=9.5)
        m.types[1].target_surface = 33.5
        m.types[1].lambda_surface = 0.1
        return (m, *tiled(m), None, None, 20)
    if name == "volume3d_o2":
        m = ref_model(32, 32, 32, seed=31)
        m.types[1].target_volume = 512.0
        return (m, *tiled(m), None, None, 4)
    if name == "volume3d_o3":
        m = ref_model(16, 16, 16, seed=32, order=3)
        m.types[1].target_volume = 512.0
        return (m, *tiled(m), None, None, 5)
    if name == "volume3d_o1":
        m = ref_model(16, 16, 16, seed=33, order=1)
        m.types[1].target_volume = 512.0
        return (m, *tiled(m), None, None, 5)
    if name == "connected":
        m = ref_model(64, 64, seed=23, contact=(0.0, 4.0, 4.0, 4.0), temperature=20.0,
                      connected=True)
        m.types[1].lambda_volume = 1.0
        return (m, *tiled(m), None, None, 20)
    if name == "connected3d":
        m = ref_model(16, 16, 16, seed=24, contact=(0.0, 4.0, 4.0, 4.0), temperature=40.0,
                      connected=True)
        m.types[1].target_volume = 512.0
        return (m, *tiled(m), None, None, 4)
    if name == "chemotaxis_ramp":
        m = ref_model(64, 64, seed=21)
        m.species = [Species("attractant", 0.0, 0.0, 0.0)]
        m.secretion = [[0.0], [0.0]]
        m.uptake = [[0.0], [0.0]]
        m.chemotaxis = [[0.0], [4.0]]
        l, t = lone_cell(m, 24, 32)
        ramp = np.tile(np.arange(64, dtype=np.float64), 64)
        return m, l, t, ramp, None, 30
    if name == "secreting":
        m = ref_model(64, 64, seed=11)
        m.species = [Species("IFNg", 0.4, 0.01, 0.0)]
        m.secretion = [[0.0], [0.2]]
        m.uptake = [[0.0], [0.01]]
        m.chemotaxis = [[0.0], [0.0]]
        return (m, *tiled(m), None, None, 30)
    if name == "length":
        m = ref_model(64, 64, seed=41, target_length=24.0, lambda_length=1.0)
        return (m, *tiled(m), None, None, 20)
    if name == "motility":
        m = ref_model(96, 96, seed=29, max_activity=20.0, lambda_activity=60.0)
        l, t = lone_cell(m, 44, 52)
        return m, l, t, None, None, 25
    if name == "drift":
        m = ref_model(64, 64, seed=30, connected=True, external=(5.0, -2.5, 0.0))
        l, t = lone_cell(m, 24, 32)
        return m, l, t, None, None, 25
    if name == "death":
        m = ref_model(64, 64, seed=5, death_rate=0.02)
        return (m, *tiled(m, side=8, nx=6, ny=6), None, None, 30)
    if name == "division":
        m = ref_model(64, 64, seed=8, division_volume=72.0)
        m.types[1].target_volume = 100.0
        m.types[1].death_rate = 0.005
        return (m, *tiled(m, side=6, nx=6, ny=6), None, None, 40)
    if name == "compaction":
        # many deaths, so the dead labels pass a quarter of the count and
        # are compacted away while cells keep dividing into fresh ids
        m = ref_model(64, 64, seed=74, division_volume=20.0, death_rate=0.06)
        m.types[1].target_volume = 24.0
        return (m, *tiled(m, side=4), None, None, 24)
    if name == "infection128":
        m, l, t = infection_sheet(128, 7)
        return m, l, t, None, None, 30
    if name == "infection3d":
        # the ensemble's volume physics on a 32^3 lattice of 8^3 cells, every
        # eighth cell infected
        m = infection_model(32, 32, 32, 64, target_volume=512.0)
        labels = tiled_labels(m, 8, 4, 4, 4)
        kinds = np.ones(64, dtype=np.uint8)
        kinds[::8] = 2
        return m, labels, default_cell_types(65, kinds), None, None, 4
    if name == "motility3d":
        m = ref_model(16, 16, 16, seed=53, max_activity=20.0, lambda_activity=60.0)
        m.types[1].target_volume = 512.0
        l, t = lone_cell(m, 4, 12)
        return m, l, t, None, None, 4
    if name == "length3d":
        m = ref_model(16, 16, 16, seed=61, order=3, target_length=14.0, lambda_length=2.0)
        m.types[1].target_volume = 512.0
        return (m, *tiled(m), None, None, 6)
    if name == "division3d":
        m = ref_model(16, 16, 16, seed=62, order=3, division_volume=72.0, death_rate=0.01)
        m.types[1].target_volume = 100.0
        return (m, *tiled(m, side=4), None, None, 10)
    if name == "nematic":
        # nematic.rs:69-113's model: random field, lambda_N 3
        m = ref_model(64, 64, seed=71, lambda_nematic=3.0)
        m.nematic = random_pairs(9, m.sites)
        return (m, *tiled(m), None, None, 20)
    if name == "nematic_division":
        m = ref_model(64, 64, seed=72, lambda_nematic=2.0, division_volume=72.0, death_rate=0.005)
        m.types[1].target_volume = 100.0
        m.nematic = random_pairs(4, m.sites)
        return (m, *tiled(m, side=6, nx=6, ny=6), None, None, 30)
    if name == "nematic3d":
        m = ref_model(16, 16, 16, seed=73, order=3, lambda_nematic=4.0)
        m.types[1].target_volume = 512.0
        m.nematic = random_pairs(6, m.sites)
        return (m, *tiled(m), None, None, 5)
    raise KeyError(name)


def random_pairs(seed, sites):
    """nematic.rs:19-24: (next_f64() - 0.5, next_f64() - 0.5) per site from
    xoshiro256++ seeded with `seed`."""
    import oracle

    rng = oracle.Xoshiro(seed)
    return np.array([[rng.next_f64() - 0.5, rng.next_f64() - 0.5] for _ in range(sites)])


PARITY_CASES = ["plain128", "surface256", "surface1024", "scattered1024", "fast_medium_flat", "fast_medium_typed",
                "fast_lone_typed", "ref64", "medium_gaps", "rect64x32",
                "von_neumann", "float_energies", "volume3d_o2", "volume3d_o3", "volume3d_o1",
                "connected", "connected3d", "chemotaxis_ramp", "secreting", "length",
                "motility", "drift", "death", "division", "compaction", "infection128", "infection3d",
                "motility3d",
                "length3d", "division3d", "nematic", "nematic_division", "nematic3d"]

"""The synthetic workloads the engine is measured and checked on.

Each builds a Model, a starting lattice and a type per label:

  plain_monolayer(side, seed, surface)  a periodic sheet of side x side sites
      tiled by 8 x 8 cells, one cell type, contact J = 10 on every unlike
      Moore bond, lambda_V = 2, V_t = 64, T = 10; with `surface`, lambda_S = 1
      and S_t = 32. 128^2 without the surface term is the CPU-runnable case;
      1024^2 to 4096^2 with it is the throughput case.
  infection_sheet(side, seed)  two cell types (epithelial 95 %, infected 5 %
      placed by seed), two f32 fields (virus D 0.1 decay 0.01, interferon D
      0.05 decay 0.02) on 4 substeps, secretion 1.0 / 0.5 by infected cells,
      uptake 0.05 by epithelial cells, chemotaxis 5 toward interferon, length
      lambda_L 1 L_t 10, infected death 0.002 per step, division at 128.
  ensemble(seed)  64 planar lattices of the infection physics with sides
      drawn by seed from {64, 128, 256, 512}, plus 16 periodic 64^3 volumes
      of 8^3 cubes (26-neighbour), each with its own seed and replica index
      and with J ~ U[5, 15], lambda_V ~ U[1, 4], T ~ U[5, 20] drawn by seed.
"""

from __future__ import annotations

import numpy as np

from .model import CellType, Model, Species, default_cell_types, tiled_labels


def plain_monolayer(side: int = 128, seed: int = 1, surface: bool = False):
    cell = CellType(target_volume=64.0, lambda_volume=2.0)
    if surface:
        cell.target_surface = 32.0
        cell.lambda_surface = 1.0
    model = Model(width=side, height=side, depth=1, contact=[0.0, 10.0, 10.0, 10.0],
                  types=[CellType(), cell], temperature=10.0, neighbour_order=2, seed=seed)
    n = side // 8
    labels = tiled_labels(model, 8, n, n, 1)
    types = default_cell_types(n * n + 1)
    return model, labels, types


def reference_model(side: int, seed: int = 1):
    """The reference's benchmark model (examples/bench.rs:7-25): contact
    medium-cell 16, cell-cell 8, V_t 64, lambda_V 2, T 10, Moore, cells of
    side 8 tiling a side x side lattice."""
    c = CellType(target_volume=64.0, lambda_volume=2.0)
    m = Model(width=side, height=side, contact=[0.0, 16.0, 16.0, 8.0], types=[CellType(), c],
              temperature=10.0, neighbour_order=2, seed=seed)
    n = side // 8
    return m, tiled_labels(m, 8, n, n, 1), default_cell_types(n * n + 1)


def infection_sheet(side: int = 512, seed: int = 7, infected_fraction: float = 0.05):
    epi = CellType(target_volume=64.0, lambda_volume=2.0, target_length=10.0, lambda_length=1.0)
    inf = CellType(target_volume=64.0, lambda_volume=2.0, target_length=10.0, lambda_length=1.0,
                   death_rate=0.002, division_volume=128.0)
    model = Model(
        width=side, height=side, depth=1,
        contact=[0.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0],
        types=[CellType(), epi, inf], temperature=10.0, neighbour_order=2, seed=seed,
        species=[Species("virus", 0.1, 0.01, 0.0, substeps=4),
                 Species("interferon", 0.05, 0.02, 0.0, substeps=4)],
        secretion=[[0.0, 0.0], [0.0, 0.0], [1.0, 0.5]],
        uptake=[[0.0, 0.0], [0.05, 0.05], [0.0, 0.0]],
        chemotaxis=[[0.0, 0.0], [0.0, 5.0], [0.0, 5.0]],
    )
    n = side // 8
    cells = n * n
    labels = tiled_labels(model, 8, n, n, 1)
    rng = np.random.default_rng(seed)
    infected = rng.choice(cells, size=int(round(infected_fraction * cells)), replace=False)
    kinds = np.ones(cells, dtype=np.uint8)
    kinds[infected] = 2
    types = default_cell_types(cells + 1, kinds)
    return model, labels, types


def infection_model(width: int, height: int, depth: int, seed: int, contact: float = 10.0,
                    lambda_volume: float = 2.0, temperature: float = 10.0,
                    target_volume: float = 64.0) -> Model:
    """The infection physics on any lattice; the division volume is twice the
    target volume (128 on the planar sheet)."""
    epi = CellType(target_volume=target_volume, lambda_volume=lambda_volume, target_length=10.0,
                   lambda_length=1.0)
    inf = CellType(target_volume=target_volume, lambda_volume=lambda_volume, target_length=10.0,
                   lambda_length=1.0, death_rate=0.002, division_volume=2.0 * target_volume)
    j = float(contact)
    return Model(
        width=width, height=height, depth=depth,
        contact=[0.0, j, j, j, j, j, j, j, j],
        types=[CellType(), epi, inf], temperature=float(temperature),
        neighbour_order=2 if depth == 1 else 3, seed=seed,
        species=[Species("virus", 0.1, 0.01, 0.0, substeps=4),
                 Species("interferon", 0.05, 0.02, 0.0, substeps=4)],
        secretion=[[0.0, 0.0], [0.0, 0.0], [1.0, 0.5]],
        uptake=[[0.0, 0.0], [0.05, 0.05], [0.0, 0.0]],
        chemotaxis=[[0.0, 0.0], [0.0, 5.0], [0.0, 5.0]],
    )


def ensemble(seed: int = 3, planar: int = 64, volumes: int = 16,
             sides=(64, 128, 256, 512), volume_side: int = 64):
    """[(model, labels, types)] of the batched ensemble: `planar` infection
    sheets of sides drawn from `sides`, then `volumes` periodic cubes of
    `volume_side`^3 tiled by 8^3 cells. Replica r has seed 1000 + r and
    replica index r; parameters are drawn from one generator seeded by `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(planar + volumes):
        three_d = r >= planar
        side = volume_side if three_d else int(rng.choice(sides))
        j = float(rng.uniform(5.0, 15.0))
        lam = float(rng.uniform(1.0, 4.0))
        temp = float(rng.uniform(5.0, 20.0))
        depth = side if three_d else 1
        m = infection_model(side, side, depth, 1000 + r, contact=j, lambda_volume=lam,
                            temperature=temp, target_volume=512.0 if three_d else 64.0)
        m.replica = r
        n = side // 8
        labels = tiled_labels(m, 8, n, n, n if three_d else 1)
        cells = n ** (3 if three_d else 2)
        kinds = np.ones(cells, dtype=np.uint8)
        kinds[rng.choice(cells, size=max(1, int(round(0.05 * cells))), replace=False)] = 2
        out.append((m, labels, default_cell_types(cells + 1, kinds)))
    return out

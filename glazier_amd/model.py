"""The model a run realises, and its flat C form for the engine's ABI.

`Model` and `CellType` mirror glazier-core's Model and CellType
(glazier-core/src/model.rs:9-76), with every term off at zero so a caller
states only what it uses. `Model.to_c()` packs one into the `ghip_model` POD
of include/glazier_hip.h. `tiled_labels` lays down the squares of
Lattice::tile_grid (glazier-core/src/lattice.rs:169-196).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

MAX_TYPES = 16
MAX_SPECIES = 4


class CCellType(ctypes.Structure):
    _fields_ = [
        ("target_volume", ctypes.c_double),
        ("lambda_volume", ctypes.c_double),
        ("target_surface", ctypes.c_double),
        ("lambda_surface", ctypes.c_double),
        ("division_volume", ctypes.c_double),
        ("death_rate", ctypes.c_double),
        ("target_length", ctypes.c_double),
        ("lambda_length", ctypes.c_double),
        ("max_activity", ctypes.c_double),
        ("lambda_activity", ctypes.c_double),
        ("external", ctypes.c_double * 3),
        ("lambda_nematic", ctypes.c_double),
        ("connected", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class CSpecies(ctypes.Structure):
    _fields_ = [
        ("diffusion", ctypes.c_double),
        ("decay", ctypes.c_double),
        ("initial", ctypes.c_double),
        ("substeps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class CModel(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("depth", ctypes.c_uint32),
        ("neighbour_order", ctypes.c_uint32),
        ("n_types", ctypes.c_uint32),
        ("n_species", ctypes.c_uint32),
        ("temperature", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("replica", ctypes.c_uint32),
        ("has_nematic", ctypes.c_uint32),
        ("contact", ctypes.c_double * (MAX_TYPES * MAX_TYPES)),
        ("types", CCellType * MAX_TYPES),
        ("species", CSpecies * MAX_SPECIES),
        ("secretion", (ctypes.c_double * MAX_SPECIES) * MAX_TYPES),
        ("uptake", (ctypes.c_double * MAX_SPECIES) * MAX_TYPES),
        ("chemotaxis", (ctypes.c_double * MAX_SPECIES) * MAX_TYPES),
    ]


@dataclass
class CellType:
    """One cell type (glazier-core/src/model.rs:9-41). Type 0 is the medium."""

    target_volume: float = 0.0
    lambda_volume: float = 0.0
    target_surface: float = 0.0
    lambda_surface: float = 0.0
    division_volume: float = 0.0
    death_rate: float = 0.0
    target_length: float = 0.0
    lambda_length: float = 0.0
    connected: bool = False
    max_activity: float = 0.0
    lambda_activity: float = 0.0
    external: Sequence[float] = (0.0, 0.0, 0.0)
    lambda_nematic: float = 0.0


@dataclass
class Species:
    """One diffusing species (glazier-core/src/field.rs:15-25). `substeps` 0
    takes the stability count (field.rs:155-160); a positive value states it."""

    name: str
    diffusion: float
    decay: float = 0.0
    initial: float = 0.0
    substeps: int = 0


@dataclass
class Model:
    """The whole model (glazier-core/src/model.rs:44-76)."""

    width: int = 1
    height: int = 1
    depth: int = 1
    contact: List[float] = field(default_factory=lambda: [0.0])
    types: List[CellType] = field(default_factory=lambda: [CellType()])
    temperature: float = 10.0
    neighbour_order: int = 2
    seed: int = 0
    species: List[Species] = field(default_factory=list)
    secretion: List[List[float]] = field(default_factory=list)  # [type][species]
    uptake: List[List[float]] = field(default_factory=list)
    chemotaxis: List[List[float]] = field(default_factory=list)
    replica: int = 0
    # Model::nematic (model.rs:70-75): (Q_xx, Q_xy) per site, shape (sites, 2)
    nematic: Optional[np.ndarray] = None

    @property
    def has_nematic(self) -> bool:
        """Model::has_nematic (model.rs:120-124): a field is present and some
        type couples to it."""
        return self.nematic is not None and len(self.nematic) > 0 and any(
            t.lambda_nematic != 0.0 for t in self.types)

    def nematic_pairs(self) -> Optional[np.ndarray]:
        """The nematic field as a contiguous f64 array of (sites, 2), or None."""
        if self.nematic is None:
            return None
        q = np.ascontiguousarray(self.nematic, dtype=np.float64).reshape(-1, 2)
        if q.shape[0] != self.sites:
            raise ValueError(f"the nematic field has {q.shape[0]} sites for a lattice of {self.sites}")
        return q

    @property
    def n_types(self) -> int:
        return len(self.types)

    @property
    def sites(self) -> int:
        return self.width * self.height * self.depth

    def shape(self):
        """Array shape of the lattice: (H, W) in a plane, (D, H, W) in a volume."""
        if self.depth > 1:
            return (self.depth, self.height, self.width)
        return (self.height, self.width)

    def validate(self) -> None:
        """Model::validate (glazier-core/src/model.rs:160-232)."""
        n = self.n_types
        if n == 0:
            raise ValueError("a model needs at least the medium type")
        if len(self.contact) != n * n:
            raise ValueError(f"contact matrix is {len(self.contact)} entries for {n} types, want {n * n}")
        for a in range(n):
            for b in range(n):
                if abs(self.contact[a * n + b] - self.contact[b * n + a]) > 1e-12:
                    raise ValueError(f"contact matrix is asymmetric at ({a}, {b})")
        if self.width == 0 or self.height == 0 or self.depth == 0:
            raise ValueError("the lattice has a zero dimension")
        if self.neighbour_order not in (1, 2, 3):
            raise ValueError(
                f"neighbour order {self.neighbour_order} is outside the range one to three")
        if self.temperature <= 0.0:
            raise ValueError("the temperature must be positive")
        s = len(self.species)
        for name, table in (("secretion", self.secretion), ("uptake", self.uptake),
                            ("chemotaxis", self.chemotaxis)):
            if table and len(table) != n:
                raise ValueError(f"{name} has {len(table)} entries for {n} types")
            for index, row in enumerate(table):
                if row and len(row) != s:
                    raise ValueError(f"type {index} states {len(row)} {name} rates for {s} species")
        if self.nematic is not None and len(self.nematic) > 0:
            self.nematic_pairs()  # model.rs:193-197

    def to_c(self) -> CModel:
        if self.n_types > MAX_TYPES:
            raise ValueError(f"at most {MAX_TYPES} cell types")
        if len(self.species) > MAX_SPECIES:
            raise ValueError(f"at most {MAX_SPECIES} species")
        c = CModel()
        c.width, c.height, c.depth = self.width, self.height, self.depth
        c.neighbour_order = self.neighbour_order
        c.n_types = self.n_types
        c.n_species = len(self.species)
        c.temperature = float(self.temperature)
        c.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        c.replica = int(self.replica)
        c.has_nematic = 1 if self.has_nematic else 0
        n = self.n_types
        for a in range(n):
            for b in range(n):
                c.contact[a * n + b] = float(self.contact[a * n + b])
        for i, t in enumerate(self.types):
            ct = c.types[i]
            ct.target_volume = t.target_volume
            ct.lambda_volume = t.lambda_volume
            ct.target_surface = t.target_surface
            ct.lambda_surface = t.lambda_surface
            ct.division_volume = t.division_volume
            ct.death_rate = t.death_rate
            ct.target_length = t.target_length
            ct.lambda_length = t.lambda_length
            ct.max_activity = t.max_activity
            ct.lambda_activity = t.lambda_activity
            for a in range(3):
                ct.external[a] = float(t.external[a])
            ct.lambda_nematic = t.lambda_nematic
            ct.connected = 1 if t.connected else 0
        for k, sp in enumerate(self.species):
            c.species[k].diffusion = sp.diffusion
            c.species[k].decay = sp.decay
            c.species[k].initial = sp.initial
            c.species[k].substeps = int(sp.substeps)
        for table, dest in ((self.secretion, c.secretion), (self.uptake, c.uptake),
                            (self.chemotaxis, c.chemotaxis)):
            for i, row in enumerate(table):
                for k, v in enumerate(row):
                    dest[i][k] = float(v)
        return c


def tiled_labels(model: Model, side: int, nx: int, ny: int, nz: int = 1) -> np.ndarray:
    """Lattice::tile_grid (glazier-core/src/lattice.rs:169-196): cubes of
    `side` laid from the origin, label = 1 + bx + nx*by (+ nx*ny*bz), the rest
    medium. Returns the flat label field, x fastest."""
    w, h, d = model.width, model.height, model.depth
    nx = min(nx, w // side)
    ny = min(ny, h // side)
    nz = 1 if d == 1 else min(nz, d // side)
    side_z = 1 if d == 1 else side
    labels = np.zeros((d, h, w), dtype=np.uint32)
    nxt = 1
    for bz in range(nz):
        for by in range(ny):
            for bx in range(nx):
                labels[bz * side_z:(bz + 1) * side_z, by * side:(by + 1) * side,
                       bx * side:(bx + 1) * side] = nxt
                nxt += 1
    return labels.reshape(-1)


def default_cell_types(n_labels: int, types: Sequence[int] | None = None) -> np.ndarray:
    """Type per label as Simulation::assemble and set_cell_types leave it
    (glazier-core/src/cpu.rs:203, 220-224): every entry 1, the medium's
    included, then `types` laid over labels 1.. in order."""
    out = np.ones(n_labels, dtype=np.uint8)
    if types is not None:
        for label, kind in enumerate(types):
            out[label + 1] = kind
    return out


class CFieldParams(ctypes.Structure):
    """`ghip_field_params` of include/glazier_hip.h."""

    _fields_ = [
        ("diffusion", ctypes.c_double * MAX_SPECIES),
        ("decay", ctypes.c_double * MAX_SPECIES),
        ("secretion", (ctypes.c_double * MAX_SPECIES) * MAX_TYPES),
        ("uptake", (ctypes.c_double * MAX_SPECIES) * MAX_TYPES),
    ]


@dataclass
class FieldParams:
    """The rates a field replay is a function of: per species diffusion and
    decay (per step, split over the substeps as the engine splits them), per
    (cell type, species) secretion and uptake (Fields::exchange,
    glazier-core/src/field.rs:133-149). Row 0 of the per-type tables is the
    medium, which exchanges nothing."""

    diffusion: np.ndarray
    decay: np.ndarray
    secretion: np.ndarray
    uptake: np.ndarray

    @property
    def n_species(self) -> int:
        return int(self.diffusion.size)

    @property
    def n_types(self) -> int:
        return int(self.secretion.shape[0])

    @staticmethod
    def from_model(model: "Model") -> "FieldParams":
        s = len(model.species)
        return FieldParams(
            diffusion=np.array([sp.diffusion for sp in model.species], dtype=np.float64),
            decay=np.array([sp.decay for sp in model.species], dtype=np.float64),
            secretion=np.array(model.secretion, dtype=np.float64).reshape(model.n_types, s),
            uptake=np.array(model.uptake, dtype=np.float64).reshape(model.n_types, s))

    def to_c(self) -> CFieldParams:
        c = CFieldParams()
        for k in range(self.n_species):
            c.diffusion[k] = float(self.diffusion[k])
            c.decay[k] = float(self.decay[k])
            for t in range(self.n_types):
                c.secretion[t][k] = float(self.secretion[t, k])
                c.uptake[t][k] = float(self.uptake[t, k])
        return c

    @staticmethod
    def from_c(c: CFieldParams, n_types: int, n_species: int) -> "FieldParams":
        return FieldParams(
            diffusion=np.array([c.diffusion[k] for k in range(n_species)], dtype=np.float64),
            decay=np.array([c.decay[k] for k in range(n_species)], dtype=np.float64),
            secretion=np.array([[c.secretion[t][k] for k in range(n_species)]
                                for t in range(n_types)], dtype=np.float64),
            uptake=np.array([[c.uptake[t][k] for k in range(n_species)]
                             for t in range(n_types)], dtype=np.float64))

    def vector(self) -> np.ndarray:
        """diffusion, decay, then the cell types' secretion and uptake rows
        (the medium row left out), as one f64 vector."""
        return np.concatenate([self.diffusion, self.decay, self.secretion[1:].ravel(),
                               self.uptake[1:].ravel()])

    @staticmethod
    def from_vector(v, n_types: int, n_species: int) -> "FieldParams":
        v = np.asarray(v, dtype=np.float64)
        s, c = n_species, (n_types - 1) * n_species
        zero = np.zeros((1, s))
        return FieldParams(
            diffusion=v[:s].copy(), decay=v[s:2 * s].copy(),
            secretion=np.vstack([zero, v[2 * s:2 * s + c].reshape(n_types - 1, s)]),
            uptake=np.vstack([zero, v[2 * s + c:2 * s + 2 * c].reshape(n_types - 1, s)]))

    def names(self) -> List[str]:
        s, t = self.n_species, self.n_types
        return ([f"diffusion[{k}]" for k in range(s)] + [f"decay[{k}]" for k in range(s)]
                + [f"secretion[{a}][{k}]" for a in range(1, t) for k in range(s)]
                + [f"uptake[{a}][{k}]" for a in range(1, t) for k in range(s)])

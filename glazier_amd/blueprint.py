"""The Blueprint description, read in Python.

The second reader of the JSON format, as python/glazier/blueprint.py is in the
reference: the same keys and defaults as glazier-core's serde reader
(glazier-core/src/blueprint.rs:15-193), the same checks (from_json,
blueprint.rs:385-447), the adhesion fold (effective_contact,
blueprint.rs:200-239) and the type assignment by fractions in label order
(initial_types, blueprint.rs:247-261). `simulation()` builds what
Blueprint::simulation builds (blueprint.rs:336-375): the model, the starting
label field and one type per label. The C++ reader behind `glazier --check`
must agree with this one on every field.

Extension: `initial.shuffle_seed` permutes the type list by a counter hash of
the seed, for an infection "uniformly placed by seed".
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

from .model import CellType, Model, Species, tiled_labels

MEDIUM = "Medium"


def hash32(x: int) -> int:
    """The engine's 32-bit finaliser (lowbias32), in Python integers."""
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def pass_key(seed: int, replica: int, step: int, colour: int) -> int:
    k = hash32((seed & 0xFFFFFFFF) ^ 0x243F6A88)
    k = hash32(k ^ ((seed >> 32) & 0xFFFFFFFF))
    k = hash32(k ^ replica)
    k = hash32(k ^ (step & 0xFFFFFFFF))
    k = hash32(k ^ ((step >> 32) & 0xFFFFFFFF))
    return hash32(k ^ colour)


def shuffled(types: List[int], seed: int) -> List[int]:
    """Fisher-Yates over `types` driven by the counter hash of `seed`."""
    out = list(types)
    key = pass_key(seed, 0, 0, 0x5EED5EED)
    for i in range(len(out) - 1, 0, -1):
        h = hash32(key ^ hash32(i))
        j = (h * (i + 1)) >> 32
        out[i], out[j] = out[j], out[i]
    return out


@dataclass
class SpeciesSpec:
    name: str
    diffusion: float
    decay: float = 0.0
    initial: float = 0.0
    substeps: int = 0


@dataclass
class TypeSpec:
    name: str
    target_volume: float
    lambda_volume: float
    secretion: Dict[str, float] = field(default_factory=dict)
    uptake: Dict[str, float] = field(default_factory=dict)
    chemotaxis: Dict[str, float] = field(default_factory=dict)
    presents: Dict[str, float] = field(default_factory=dict)
    target_surface: float = 0.0
    lambda_surface: float = 0.0
    division_volume: float = 0.0
    death_rate: float = 0.0
    target_length: float = 0.0
    lambda_length: float = 0.0
    connected: bool = False
    max_activity: float = 0.0
    lambda_activity: float = 0.0
    external: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    lambda_nematic: float = 0.0


_TYPE_KEYS = set(TypeSpec.__dataclass_fields__)
_SPECIES_KEYS = set(SpeciesSpec.__dataclass_fields__)


@dataclass
class Blueprint:
    name: str
    width: int
    height: int
    types: List[TypeSpec]
    contact: List[float]
    temperature: float
    neighbour_order: int
    seed: int
    steps: int
    dump_every: int
    initial: dict
    units: dict
    depth: int = 1
    fields: List[SpeciesSpec] = field(default_factory=list)
    adhesion: Optional[dict] = None
    nematic_field: Optional[str] = None
    base: Path = Path(".")

    @property
    def dimensions(self) -> int:
        """Two for a plane, three for a volume (python/glazier/blueprint.py:74-77)."""
        return 3 if self.depth > 1 else 2

    @property
    def type_names(self) -> List[str]:
        return [MEDIUM] + [t.name for t in self.types]

    @property
    def species_names(self) -> List[str]:
        return [s.name for s in self.fields]

    @staticmethod
    def from_json(text: str, base: Path | str = ".") -> "Blueprint":
        raw = json.loads(text)
        for key in ("name", "width", "height", "types", "contact", "temperature",
                    "neighbour_order", "seed", "steps", "dump_every", "initial", "units"):
            if key not in raw:
                raise ValueError(f"missing field `{key}`")
        types = [TypeSpec(**{k: v for k, v in t.items() if k in _TYPE_KEYS}) for t in raw["types"]]
        species = [SpeciesSpec(**{k: v for k, v in f.items() if k in _SPECIES_KEYS})
                   for f in raw.get("fields", [])]
        bp = Blueprint(name=raw["name"], width=int(raw["width"]), height=int(raw["height"]),
                       types=types, contact=[float(c) for c in raw["contact"]],
                       temperature=float(raw["temperature"]),
                       neighbour_order=int(raw["neighbour_order"]), seed=int(raw["seed"]),
                       steps=int(raw["steps"]), dump_every=int(raw["dump_every"]),
                       initial=dict(raw["initial"]), units=dict(raw["units"]),
                       depth=int(raw.get("depth", 1)), fields=species,
                       adhesion=raw.get("adhesion"), nematic_field=raw.get("nematic_field"),
                       base=Path(base))
        bp._check()
        return bp

    @staticmethod
    def load(path: Path | str) -> "Blueprint":
        path = Path(path)
        return Blueprint.from_json(path.read_text(), base=path.parent)

    def _check(self) -> None:
        """The checks of Blueprint::from_json (blueprint.rs:386-446)."""
        names = [t.name for t in self.types]
        for name in self.initial.get("fractions", {}):
            if name not in names:
                raise ValueError(f"the initial condition names the type {name}, which the "
                                 "description does not state")
        if self.adhesion is not None:
            n = len(self.adhesion["molecules"])
            if len(self.adhesion["binding"]) != n * n:
                raise ValueError(f"the binding matrix is {len(self.adhesion['binding'])} entries "
                                 f"for {n} molecules, want {n * n}")
            for t in self.types:
                for name in t.presents:
                    if name not in self.adhesion["molecules"]:
                        raise ValueError(f"type {t.name} presents {name}, which is not a molecule "
                                         "this description declares")
        else:
            for t in self.types:
                if t.presents:
                    raise ValueError(f"type {t.name} presents adhesion molecules, and the "
                                     "description declares none")
        known = self.species_names
        for t in self.types:
            for what, mapping in (("secretion", t.secretion), ("uptake", t.uptake),
                                  ("chemotaxis", t.chemotaxis)):
                for name in mapping:
                    if name not in known:
                        raise ValueError(f"type {t.name} states {what} of {name}, which is not a "
                                         "field this description declares")
        self.model().validate()

    def effective_contact(self) -> List[float]:
        contact = list(self.contact)
        if self.adhesion is None:
            return contact
        molecules = self.adhesion["molecules"]
        binding = self.adhesion["binding"]
        n = len(molecules)
        presented = [[0.0] * n] + [[t.presents.get(m, 0.0) for m in molecules] for t in self.types]
        nt = len(presented)
        for a in range(nt):
            for b in range(nt):
                bound = 0.0
                for i in range(n):
                    for j in range(n):
                        bound += binding[i * n + j] * presented[a][i] * presented[b][j]
                contact[a * nt + b] -= bound
        return contact

    def initial_types(self, cells: int) -> List[int]:
        out = [1] * cells
        nxt = 0
        fractions = self.initial.get("fractions", {})
        for index, spec in enumerate(self.types):
            if spec.name not in fractions:
                continue
            x = fractions[spec.name] * cells
            share = int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))
            share = min(share, cells - nxt)
            for k in range(nxt, nxt + share):
                out[k] = index + 1
            nxt += share
        if "shuffle_seed" in self.initial:
            out = shuffled(out, int(self.initial["shuffle_seed"]))
        return out

    def model(self) -> Model:
        species = [Species(f.name, f.diffusion, f.decay, f.initial, f.substeps) for f in self.fields]
        names = self.species_names

        def rates(mapping):
            return [float(mapping.get(n, 0.0)) for n in names]

        types = [CellType()] + [CellType(
            target_volume=t.target_volume, lambda_volume=t.lambda_volume,
            target_surface=t.target_surface, lambda_surface=t.lambda_surface,
            division_volume=t.division_volume, death_rate=t.death_rate,
            target_length=t.target_length, lambda_length=t.lambda_length,
            connected=bool(t.connected), max_activity=t.max_activity,
            lambda_activity=t.lambda_activity, external=tuple(t.external),
            lambda_nematic=t.lambda_nematic) for t in self.types]
        zeros = [0.0] * len(names)
        return Model(width=self.width, height=self.height, depth=self.depth,
                     contact=self.effective_contact(), types=types, temperature=self.temperature,
                     neighbour_order=self.neighbour_order, seed=self.seed, species=species,
                     secretion=[zeros] + [rates(t.secretion) for t in self.types],
                     uptake=[zeros] + [rates(t.uptake) for t in self.types],
                     chemotaxis=[zeros] + [rates(t.chemotaxis) for t in self.types],
                     nematic=self.nematic())

    def nematic(self) -> Optional[np.ndarray]:
        """The nematic field the description names (blueprint.rs:347-358): a .npy
        of shape (height, width, 2) or (depth, height, width, 2), f8 or f4,
        resolved against the description's directory, as (sites, 2) f64."""
        if self.nematic_field is None:
            return None
        path = Path(self.nematic_field)
        if not path.is_absolute():
            path = self.base / path
        raw = np.load(path, allow_pickle=False)
        if raw.ndim < 1 or raw.shape[-1] != 2:
            raise ValueError(f"{path}: the last axis is {raw.shape[-1:]}, want 2")
        if raw.dtype not in (np.float64, np.float32):
            raise ValueError(f"{path}: a field of dtype {raw.dtype} is not read")
        pairs = raw.astype(np.float64).reshape(-1, 2)
        sites = self.width * self.height * self.depth
        if pairs.shape[0] != sites:
            raise ValueError(f"{path}: shape {raw.shape} holds {pairs.shape[0]} sites for a lattice of {sites}")
        return pairs

    def simulation(self):
        """(model, labels, cell types) as Blueprint::simulation builds them."""
        model = self.model()
        if self.initial.get("labels"):
            labels = np.load(self.base / self.initial["labels"]).astype(np.uint32).reshape(-1)
            cells = int(labels.max())
        else:
            labels = tiled_labels(model, self.initial["side"], self.initial["nx"],
                                  self.initial["ny"], self.initial.get("nz", 1))
            cells = int(labels.max())
        types = np.ones(cells + 1, dtype=np.uint8)
        types[1:] = self.initial_types(cells)
        return model, labels, types

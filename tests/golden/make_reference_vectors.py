"""Writes the reader and measurement vectors of the reference's pure-Python
modules, for the tests that pin this build's own readers against them.

Run in the build container only, where the reference tree is present:

    python tests/golden/make_reference_vectors.py

It loads python/glazier/blueprint.py and python/glazier/director.py of the
reference by file path (not through the package, so nothing else of the
package is imported) and writes data only:

* reference_readings.json: for every blueprint under tests/golden/blueprints,
  what the reference's Blueprint.load reads (every dataclass field, the
  effective contact matrix, type and species names, dimensions, physical
  domain and duration);
* director_vectors.npz: a few fixed label fields and, for each, what the
  reference's measurements return (cell ellipses, the painted Q field, its
  Gaussian coarse-graining, director angle, scalar order, global order and
  the disclination list).

The tests read these files; they never import the reference. Not collected by
pytest (the name does not start with test_) and not sent to a GPU run.
"""

from __future__ import annotations

import dataclasses
import importlib.util
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REFERENCE = Path("/root/reference/python/glazier")
BLUEPRINTS = HERE / "blueprints"
SIGMA = 4.0


def load_module(name: str, path: Path):
    spec = importlib.util.spec_from_file_location(name, path)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module  # dataclasses resolve their module by name
    spec.loader.exec_module(module)
    return module


def reading(bp) -> dict:
    out = dataclasses.asdict(bp)
    out["effective_contact"] = bp.effective_contact()
    out["type_names"] = bp.type_names
    out["species_names"] = bp.species_names
    out["dimensions"] = bp.dimensions
    out["domain_microns"] = list(bp.domain_microns)
    out["duration_minutes"] = bp.duration_minutes
    return out


def label_fields() -> dict:
    """Fixed label fields: cells with ragged borders, a brick of elongated
    cells with medium gaps and a sliver under four sites, and a volume."""
    rng = np.random.default_rng(20251204)
    fields = {}

    side = 64
    y, x = np.mgrid[0:side, 0:side]
    tiles = (1 + x // 8 + 8 * (y // 8)).astype(np.uint32)
    for _ in range(3000):
        i, j = rng.integers(0, side, size=2)
        di, dj = rng.integers(-1, 2, size=2)
        tiles[i, j] = tiles[(i + di) % side, (j + dj) % side]
    fields["ragged"] = tiles

    brick = np.zeros((side, side), dtype=np.uint32)
    label = 1
    for row in range(side // 4):
        shift = 8 if row % 2 else 0
        for col in range(side // 16):
            x0 = (col * 16 + shift) % side
            for dy in range(4):
                for dx in range(15):  # one medium column between bricks
                    brick[row * 4 + dy, (x0 + dx) % side] = label
            label += 1
    for _ in range(600):
        i, j = rng.integers(0, side, size=2)
        di, dj = rng.integers(-1, 2, size=2)
        brick[i, j] = brick[(i + di) % side, (j + dj) % side]
    brick[10, 10] = label  # a cell of one site: measured as too small
    brick[30, 31] = label + 1
    brick[30, 32] = label + 1
    fields["brick"] = brick

    vol = np.zeros((8, 32, 32), dtype=np.uint32)
    zz, yy, xx = np.mgrid[0:8, 0:32, 0:32]
    vol[:] = 1 + xx // 8 + 4 * (yy // 8) + 16 * (zz // 4)
    for _ in range(1500):
        k, i, j = rng.integers(0, 8), rng.integers(0, 32), rng.integers(0, 32)
        dk, di, dj = rng.integers(-1, 2, size=3)
        vol[k, i, j] = vol[(k + dk) % 8, (i + di) % 32, (j + dj) % 32]
    fields["volume"] = vol
    return fields


def main() -> None:
    blueprint = load_module("reference_blueprint", REFERENCE / "blueprint.py")
    director = load_module("reference_director", REFERENCE / "director.py")

    readings = {}
    for path in sorted(BLUEPRINTS.glob("*.json")):
        readings[path.name] = reading(blueprint.Blueprint.load(path))
    (HERE / "reference_readings.json").write_text(json.dumps(readings, indent=1, sort_keys=True) + "\n")

    arrays = {}
    for name, labels in label_fields().items():
        arrays[f"{name}__labels"] = labels
        plane = director.mid_plane(labels)
        arrays[f"{name}__mid_plane"] = plane
        shapes = director.cell_shapes(plane)
        # label, area, com x, com y, theta, eccentricity per measured cell
        arrays[f"{name}__shapes"] = np.array(
            [[s.label, s.area, s.com[0], s.com[1], s.theta, s.eccentricity] for s in shapes],
            dtype=np.float64).reshape(-1, 6)
        q = director.q_field(plane, shapes)
        arrays[f"{name}__q"] = q
        qc = director.coarse_grain(q, SIGMA)
        arrays[f"{name}__q_coarse"] = qc
        theta = director.director_angle(qc)
        arrays[f"{name}__theta"] = theta
        arrays[f"{name}__order"] = director.scalar_order(qc)
        arrays[f"{name}__global_order"] = np.array(director.global_order(shapes), dtype=np.float64)
        arrays[f"{name}__defects"] = np.array(director.defect_charges(theta), dtype=np.float64).reshape(-1, 3)
    np.savez_compressed(HERE / "director_vectors.npz", **arrays)
    print(f"wrote {len(readings)} readings and {len(arrays)} arrays")


if __name__ == "__main__":
    main()

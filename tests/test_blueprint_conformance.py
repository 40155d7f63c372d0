"""Both readers of a Blueprint have to agree about what it says.

Restates the reference's conformance test (tests/test_blueprint_conformance.py:
45-123): the Python reader here (glazier_amd.blueprint) against what the C++
reader of the `glazier` binary prints with `--check`, over every blueprint the
reference ships (copied as fixtures into tests/golden/blueprints).
"""

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from glazier_amd.blueprint import Blueprint

ROOT = Path(__file__).resolve().parents[1]
BINARY = ROOT / "glazier_amd" / "glazier"
BLUEPRINTS = sorted((ROOT / "tests" / "golden" / "blueprints").glob("*.json"))


def binary_reading(path):
    if not BINARY.exists():
        pytest.fail(f"{BINARY} is not built; run __graft_entry__.build()")
    done = subprocess.run([str(BINARY), "--model", str(path), "--check"], check=True,
                          capture_output=True, text=True)
    return json.loads(done.stdout)


def assert_agrees(bp, rust):
    assert rust["name"] == bp.name
    assert rust["width"] == bp.width
    assert rust["height"] == bp.height
    assert rust["n_types"] == len(bp.type_names)
    assert list(rust["contact"]) == pytest.approx(bp.effective_contact())
    assert rust["temperature"] == bp.temperature
    assert rust["neighbour_order"] == bp.neighbour_order
    assert rust["seed"] == bp.seed
    assert rust["steps"] == bp.steps
    assert list(rust["initial"])[:3] == [bp.initial["side"], bp.initial["nx"], bp.initial["ny"]]
    assert rust["initial"][3] == bp.initial.get("nz", 1)
    assert rust["micron_per_site"] == bp.units["micron_per_site"]
    assert rust["minute_per_step"] == bp.units["minute_per_step"]
    assert rust["target_volume"][0] == 0.0
    assert rust["lambda_volume"][0] == 0.0
    for index, spec in enumerate(bp.types, start=1):
        assert rust["target_volume"][index] == spec.target_volume
        assert rust["lambda_volume"][index] == spec.lambda_volume
        assert rust["target_surface"][index] == spec.target_surface
        assert rust["lambda_surface"][index] == spec.lambda_surface
        assert rust["division_volume"][index] == spec.division_volume
        assert rust["death_rate"][index] == spec.death_rate
        assert rust["target_length"][index] == spec.target_length
        assert rust["lambda_length"][index] == spec.lambda_length
        assert bool(rust["connected"][index]) == spec.connected
        assert rust["lambda_nematic"][index] == spec.lambda_nematic
        assert rust["max_activity"][index] == spec.max_activity
        assert rust["lambda_activity"][index] == spec.lambda_activity
        assert list(rust["external"][index]) == list(spec.external)
    assert list(rust["species"]) == bp.species_names
    for index, species in enumerate(bp.fields):
        assert rust["diffusion"][index] == species.diffusion
        assert rust["decay"][index] == species.decay
    for type_index, spec in enumerate(bp.types, start=1):
        for species_index, name in enumerate(bp.species_names):
            assert rust["secretion"][type_index][species_index] == spec.secretion.get(name, 0.0)
            assert rust["uptake"][type_index][species_index] == spec.uptake.get(name, 0.0)
            assert rust["chemotaxis"][type_index][species_index] == spec.chemotaxis.get(name, 0.0)
    assert rust["nematic_field"] == bp.nematic_field
    assert rust["initial_labels"] == bp.initial.get("labels")


@pytest.mark.parametrize("path", BLUEPRINTS, ids=lambda p: p.stem)
def test_the_binary_agrees_with_python(path):
    assert_agrees(Blueprint.load(path), binary_reading(path))


def test_every_blueprint_in_the_tree_is_covered():
    assert len(BLUEPRINTS) == 13


def test_adhesion_folds_into_contact():
    # adhesion.rs:36-58
    text = (ROOT / "tests" / "golden" / "blueprints" / "adhesion.json").read_text()
    bp = Blueprint.from_json(text)
    n = len(bp.type_names)
    c = bp.effective_contact()
    assert c != bp.contact
    assert c[0 * n + 1] == bp.contact[0 * n + 1]  # the medium presents nothing


@pytest.mark.parametrize("bad,message", [
    ({"initial": {"side": 8, "nx": 2, "ny": 2, "fractions": {"Nope": 0.5}}}, "names the type Nope"),
    ({"types": [{"name": "Cell", "target_volume": 64.0, "lambda_volume": 2.0,
                 "secretion": {"ghost": 1.0}}]}, "secretion of ghost"),
    ({"types": [{"name": "Cell", "target_volume": 64.0, "lambda_volume": 2.0,
                 "presents": {"cad": 1.0}}]}, "declares none"),
    ({"contact": [0.0, 16.0, 15.0, 8.0]}, "asymmetric"),
    ({"neighbour_order": 4}, "outside the range"),
])
def test_both_readers_refuse_the_same_descriptions(tmp_path, bad, message):
    base = json.loads((ROOT / "tests" / "golden" / "blueprints" / "monolayer.json").read_text())
    base.update(bad)
    path = tmp_path / "bad.json"
    path.write_text(json.dumps(base))
    with pytest.raises(ValueError) as err:
        Blueprint.load(path)
    assert message in str(err.value)
    done = subprocess.run([str(BINARY), "--model", str(path), "--check"], capture_output=True,
                          text=True)
    assert done.returncode == 1
    assert message in done.stderr


def test_unknown_flags_and_engines_print_usage(tmp_path):
    # main.rs:9-21, 147-150: usage and exit code 2
    done = subprocess.run([str(BINARY), "--frobnicate"], capture_output=True, text=True)
    assert done.returncode == 2 and "usage" in done.stderr
    bp = ROOT / "tests" / "golden" / "blueprints" / "monolayer.json"
    done = subprocess.run([str(BINARY), "--model", str(bp), "--out", str(tmp_path), "--engine",
                           "quantum"], capture_output=True, text=True)
    assert done.returncode == 2


def test_initial_types_follow_label_order_and_shuffle_by_seed():
    # blueprint.rs:247-261: fractions handed out in label order
    text = (ROOT / "tests" / "golden" / "blueprints" / "infection.json").read_text()
    bp = Blueprint.from_json(text)
    t = bp.initial_types(256)
    infected = int(round(0.02 * 256))
    assert t[:infected] == [2] * infected and all(x == 1 for x in t[infected:])
    raw = json.loads(text)
    raw["initial"]["shuffle_seed"] = 7
    sh = Blueprint.from_json(json.dumps(raw)).initial_types(256)
    assert sorted(sh) == sorted(t) and sh != t

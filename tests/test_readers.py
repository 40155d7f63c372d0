"""Blueprint readers pinned to the reference's own reading of every shipped
description.

`tests/golden/reference_readings.json` holds what the reference's Python
reader (python/glazier/blueprint.py) makes of each blueprint under
tests/golden/blueprints, written by tests/golden/make_reference_vectors.py.
Three readers of this build are held to it:

* `glazier_amd.blueprint.Blueprint`, field by field;
* the C++ reader, as the `glazier --check` JSON (glazier/src/main.rs:60-96
  names the keys) and as `ghip_blueprint_parse` (the `parse` of the session
  surface, which adds `depth`).

The expected `--check` object is built from the reference reading by a table
of key -> rule, so a key the binary prints and the table lacks, or the
reverse, fails as well as a wrong value.
"""

import json
import subprocess
from pathlib import Path

import pytest

from glazier_amd.blueprint import Blueprint

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
BINARY = ROOT / "glazier_amd" / "glazier"
READINGS = json.loads((GOLDEN / "reference_readings.json").read_text())
NAMES = sorted(READINGS)

# Per-type keys of --check: the medium leads with the defaults of an
# unconstrained type (model.rs:9-41), then the types in description order.
PER_TYPE = {
    "target_volume": 0.0, "lambda_volume": 0.0, "target_surface": 0.0, "lambda_surface": 0.0,
    "target_length": 0.0, "lambda_length": 0.0, "connected": False, "max_activity": 0.0,
    "lambda_activity": 0.0, "external": [0.0, 0.0, 0.0], "lambda_nematic": 0.0,
    "division_volume": 0.0, "death_rate": 0.0,
}
PER_EXCHANGE = ("secretion", "uptake", "chemotaxis")


def _exchange(r, key):
    medium = [0.0] * len(r["species_names"])
    return [medium] + [[t[key].get(s, 0.0) for s in r["species_names"]] for t in r["types"]]


RULES = {
    "name": lambda r: r["name"],
    "width": lambda r: r["width"],
    "height": lambda r: r["height"],
    "n_types": lambda r: len(r["type_names"]),
    "contact": lambda r: r["effective_contact"],
    "temperature": lambda r: r["temperature"],
    "neighbour_order": lambda r: r["neighbour_order"],
    "seed": lambda r: r["seed"],
    "steps": lambda r: r["steps"],
    "initial": lambda r: [r["initial"][k] for k in ("side", "nx", "ny")] + [r["initial"].get("nz", 1)],
    "micron_per_site": lambda r: r["units"]["micron_per_site"],
    "minute_per_step": lambda r: r["units"]["minute_per_step"],
    "species": lambda r: r["species_names"],
    "diffusion": lambda r: [f["diffusion"] for f in r["fields"]],
    "decay": lambda r: [f["decay"] for f in r["fields"]],
    "nematic_field": lambda r: r["nematic_field"],
    "initial_labels": lambda r: r["initial"].get("labels"),
}
for _key, _medium in PER_TYPE.items():
    RULES[_key] = (lambda k, m: lambda r: [m] + [t[k] for t in r["types"]])(_key, _medium)
for _key in PER_EXCHANGE:
    RULES[_key] = (lambda k: lambda r: _exchange(r, k))(_key)


def expected_check(name):
    return {key: rule(READINGS[name]) for key, rule in RULES.items()}


def compare(got, want, where):
    """Exact equality, except the folded contact matrix (float arithmetic in
    two languages), which agrees to 1e-12."""
    assert sorted(got) == sorted(want), f"{where}: keys differ"
    for key in want:
        if key == "contact":
            assert got[key] == pytest.approx(want[key], rel=1e-12, abs=1e-12), f"{where}: {key}"
        elif key == "connected":
            assert [bool(v) for v in got[key]] == want[key], f"{where}: {key}"
        else:
            assert got[key] == want[key], f"{where}: {key}"


def check_output(path):
    if not BINARY.exists():
        pytest.fail(f"{BINARY} is not built; run __graft_entry__.build()")
    done = subprocess.run([str(BINARY), "--model", str(path), "--check"], capture_output=True,
                          text=True, check=True)
    return json.loads(done.stdout)


def test_the_fixture_covers_every_shipped_blueprint():
    shipped = sorted(p.name for p in (GOLDEN / "blueprints").glob("*.json"))
    assert shipped == NAMES and len(NAMES) == 13


@pytest.mark.parametrize("name", NAMES)
def test_check_output_matches_the_reference_reading(name):
    compare(check_output(GOLDEN / "blueprints" / name), expected_check(name), f"--check {name}")


@pytest.mark.parametrize("name", NAMES)
def test_library_parse_matches_the_reference_reading(name):
    from glazier_amd import native

    got = native.parse((GOLDEN / "blueprints" / name).read_text())
    depth = got.pop("depth")
    assert depth == READINGS[name]["depth"]
    compare(got, expected_check(name), f"parse {name}")


SCALARS = ("name", "width", "height", "depth", "temperature", "neighbour_order", "seed", "steps",
           "dump_every", "initial", "units", "contact", "adhesion", "nematic_field")
DERIVED = {
    "effective_contact": lambda bp: bp.effective_contact(),
    "type_names": lambda bp: bp.type_names,
    "species_names": lambda bp: bp.species_names,
    "dimensions": lambda bp: bp.dimensions,
}


@pytest.mark.parametrize("name", NAMES)
def test_python_reader_matches_the_reference_reading(name):
    ref = READINGS[name]
    bp = Blueprint.load(GOLDEN / "blueprints" / name)
    for key in SCALARS:
        assert getattr(bp, key) == ref[key], key
    for key, read in DERIVED.items():
        assert read(bp) == ref[key], key
    assert len(bp.types) == len(ref["types"])
    for spec, want in zip(bp.types, ref["types"]):
        for key, value in want.items():
            got = getattr(spec, key)
            assert (list(got) if key == "external" else got) == value, f"{want['name']}.{key}"
    assert [(f.name, f.diffusion, f.decay, f.initial) for f in bp.fields] == [
        (f["name"], f["diffusion"], f["decay"], f["initial"]) for f in ref["fields"]]


def test_adhesion_folds_into_contact():
    # blueprint.rs:200-239: what two types' molecules bind comes off their
    # contact energy; the medium presents nothing
    bp = Blueprint.load(GOLDEN / "blueprints" / "adhesion.json")
    n = len(bp.type_names)
    c = bp.effective_contact()
    assert c != bp.contact
    assert c[:n] == bp.contact[:n]
    assert c[n + 1] < bp.contact[n + 1]


# (change to monolayer.json, message both readers give)
REFUSALS = [
    ({"initial": {"side": 8, "nx": 2, "ny": 2, "fractions": {"Nope": 0.5}}}, "names the type Nope"),
    ({"types": [{"name": "Cell", "target_volume": 64.0, "lambda_volume": 2.0,
                 "secretion": {"ghost": 1.0}}]}, "secretion of ghost"),
    ({"types": [{"name": "Cell", "target_volume": 64.0, "lambda_volume": 2.0,
                 "presents": {"cad": 1.0}}]}, "declares none"),
    ({"contact": [0.0, 16.0, 15.0, 8.0]}, "asymmetric"),
    ({"neighbour_order": 4}, "outside the range"),
]


@pytest.mark.parametrize("change,message", REFUSALS, ids=[m for _, m in REFUSALS])
def test_both_readers_refuse_the_same_descriptions(tmp_path, change, message):
    text = json.loads((GOLDEN / "blueprints" / "monolayer.json").read_text())
    text.update(change)
    path = tmp_path / "bad.json"
    path.write_text(json.dumps(text))
    with pytest.raises(ValueError, match=message):
        Blueprint.load(path)
    done = subprocess.run([str(BINARY), "--model", str(path), "--check"], capture_output=True,
                          text=True)
    assert (done.returncode, message in done.stderr) == (1, True)


def test_unknown_flags_and_engines_print_usage(tmp_path):
    # main.rs:9-21, 147-150: usage and exit code 2
    done = subprocess.run([str(BINARY), "--frobnicate"], capture_output=True, text=True)
    assert done.returncode == 2 and "usage" in done.stderr
    done = subprocess.run([str(BINARY), "--model", str(GOLDEN / "blueprints" / "monolayer.json"),
                           "--out", str(tmp_path), "--engine", "quantum"], capture_output=True,
                          text=True)
    assert done.returncode == 2


def test_initial_types_follow_label_order_and_shuffle_by_seed():
    # blueprint.rs:247-261: fractions handed out in label order
    text = (GOLDEN / "blueprints" / "infection.json").read_text()
    t = Blueprint.from_json(text).initial_types(256)
    infected = int(round(0.02 * 256))
    assert t[:infected] == [2] * infected and set(t[infected:]) == {1}
    raw = json.loads(text)
    raw["initial"]["shuffle_seed"] = 7
    shuffled = Blueprint.from_json(json.dumps(raw)).initial_types(256)
    assert sorted(shuffled) == sorted(t) and shuffled != t

"""The `glazier` binary on the HIP engine: outputs and their invariants.

As the reference's native smoke test (tests/test_native_engine.py:27-48) and
the summary keys of glazier/src/main.rs:167-185.
"""

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BINARY = ROOT / "glazier_amd" / "glazier"
GOLDEN = ROOT / "tests" / "golden" / "blueprints"


def run(tmp_path, name, steps=20, trace=False):
    raw = json.loads((GOLDEN / name).read_text())
    raw["steps"] = steps
    raw["dump_every"] = steps // 2
    path = tmp_path / name
    path.write_text(json.dumps(raw))
    out = tmp_path / "out"
    cmd = [str(BINARY), "--model", str(path), "--out", str(out), "--engine", "hip"]
    if trace:
        cmd.append("--trace")
    print(f"[cli] {name} {steps} steps", flush=True)
    done = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert done.returncode == 0, done.stderr
    return raw, out, done.stdout


def test_a_run_writes_the_reference_outputs(tmp_path):
    raw, out, stdout = run(tmp_path, "monolayer.json", 20, trace=True)
    summary = json.loads((out / "summary.json").read_text())
    for key in ("name", "engine", "steps", "seconds", "setup_seconds", "overall_seconds",
                "n_cells", "n_labels", "fields", "mean_volume", "width", "height",
                "micron_per_site", "minute_per_step"):
        assert key in summary
    labels = np.load(out / "labels_00020.npy")
    assert labels.dtype == np.uint32 and labels.shape == (128, 128)
    assert np.unique(labels[labels > 0]).size == 256
    assert (out / "labels_00010.npy").exists()
    assert summary["n_cells"] == 256 and abs(summary["mean_volume"] - 64.0) < 1e-9
    rows = (out / "trace.csv").read_text().strip().splitlines()
    assert rows[0].startswith("step,cells,mean_volume") and len(rows) == 21
    assert stdout.startswith("hip: 20 steps in")


def test_an_infection_run_keeps_its_field_and_counts(tmp_path):
    raw, out, _ = run(tmp_path, "infection.json", 30)
    summary = json.loads((out / "summary.json").read_text())
    assert set(summary["fields"]) == {"virus", "interferon"}
    assert summary["fields"]["virus"] > 0.0
    assert (out / "field_virus_00015.npy").exists()
    labels = np.load(out / "labels_00030.npy")
    assert summary["n_cells"] == np.unique(labels[labels > 0]).size


def test_a_nematic_blueprint_runs_with_its_field(tmp_path):
    # monolayer-nematic.json names its field beside itself (blueprint.rs:347-358)
    raw = json.loads((GOLDEN / "monolayer-nematic.json").read_text())
    raw["steps"] = 40
    raw["dump_every"] = 20
    raw["nematic_field"] = str(GOLDEN / raw["nematic_field"])
    path = tmp_path / "monolayer-nematic.json"
    path.write_text(json.dumps(raw))
    out = tmp_path / "out"
    print("[cli] monolayer-nematic.json 40 steps", flush=True)
    done = subprocess.run([str(BINARY), "--model", str(path), "--out", str(out), "--engine", "hip"],
                          capture_output=True, text=True, timeout=120)
    assert done.returncode == 0, done.stderr
    summary = json.loads((out / "summary.json").read_text())
    assert summary["n_cells"] == 256 and abs(summary["mean_volume"] - 64.0) < 1e-9
    assert np.load(out / "labels_00040.npy").shape == (128, 128)

"""Benchmark of the hot path: the checkerboard Metropolis sweep at 4096^2.

Workload: the plain monolayer at 4096 x 4096 sites (262 144 cells of side 8,
u32 labels), contact J = 10, lambda_V = 2, V_t = 64, lambda_S = 1, S_t = 32,
T = 10, Moore neighbourhood. One "step" is one Monte Carlo step (four colour
passes, 4096^2 copy attempts) of one lattice; the timed steps go round-robin
over four lattices seeded 1-4, so --steps 400 runs each seed for 100 MCS.
Inputs are resident in HBM before the timed region starts.

Prints one JSON line with site-attempts/s, MCS/s, the roofline of the sweep
kernel (algorithmic bytes / measured launch time against 8 TB/s) and a CPU
baseline (the serial oracle on this host, one thread, on a bounded sample).

The same line carries "ensemble": the batched-versus-looped comparison on
the uneven ensemble (64 infection sheets of 64^2..512^2 plus 16 volumes of
64^3, parameters drawn by seed 3), each lattice run for --ensemble-mcs steps
once as one batch and once lattice by lattice; replicas/s is lattices
finished per second of wall time, setup excluded. "configs" has the other
workloads (config 0, config 1 at 1024^2 and 2048^2, config 2, and the
reference's own benchmark model beside its published device rates). And
"adjoint": the
gradient of the 256^2 field fit (forward and backward times, agreement with
the f64 CPU adjoint, and 50 Adam steps on the perturbed rates).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(side: int, seconds: float):
    """The serial oracle timed on this host: one thread, whole MCS of the same
    model at `side`^2 until about `seconds` have run."""
    import oracle
    from glazier_amd.workloads import plain_monolayer

    model, labels, types = plain_monolayer(side, seed=1, surface=True)
    sim = oracle.SerialSimulation(model, labels, types)
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    rate = steps * side * side / el
    return {"value": rate, "unit": "site-attempts/s", "cores": 1, "kind": "port",
            "sample": f"serial f64 oracle, {side}^2 plain monolayer with surface term, seed 1, "
                      f"{steps} MCS in {el:.1f} s on one thread of {os.cpu_count()} "
                      f"({cpu_model()})"}


def ensemble_bench(mcs: int):
    """Batched against looped on the uneven ensemble: the same lattices from
    the same starting state, `mcs` steps each, timed from the first launch to
    the last completion (device allocation and upload excluded)."""
    import torch

    from glazier_amd import HipBatch, HipSimulation
    from glazier_amd.workloads import ensemble

    members = ensemble(seed=3)
    sites = sum(int(l.size) for _, l, _ in members)
    # load every kernel variant once, alone and batched, outside the timing
    warm = [HipSimulation(*members[i]) for i in (0, len(members) - 1)]
    for w in warm:
        w.step(1)
    wb = HipBatch(warm)
    wb.set_tiles(0)
    wb.step(1)
    wb.close()
    for w in warm:
        w.close()

    sims = [HipSimulation(m, l, t) for m, l, t in members]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.step(mcs)
    torch.cuda.synchronize()
    looped = time.perf_counter() - t0
    looped_labels = [s.labels() for s in sims]
    for s in sims:
        s.close()

    sims = [HipSimulation(m, l, t) for m, l, t in members]
    batch = HipBatch(sims)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.step(mcs)
    torch.cuda.synchronize()
    batched = time.perf_counter() - t0
    same = all((s.labels() == want).all() for s, want in zip(sims, looped_labels))
    # algorithmic bytes of the batched run (the per-unit figures the single
    # lattice's rooflines use): every colour pass reads the labels and the
    # chemotactic fields, an accepted copy writes a label, a field step reads
    # and writes each species and reads the owners, the moments read labels
    accepted = 0
    algo = 0.0
    for (m, l, t), sim in zip(members, sims):
        acc = sim.counters()[1]
        accepted += acc
        colours = 8 if m.depth > 1 else 4
        chem = sum(1 for k in range(len(m.species)) if any(row[k] != 0.0 for row in m.chemotaxis))
        per_mcs = colours * (4.0 + 4.0 * chem) + 8.0 * len(m.species) + (4.0 if m.species else 0.0) + 4.0
        algo += per_mcs * l.size * mcs + 4.0 * acc
    # the captured step alone, after tuning
    t0 = time.perf_counter()
    batch.step(10)
    torch.cuda.synchronize()
    graph_step = (time.perf_counter() - t0) / 10
    info = batch.describe()
    batch.close()
    for s in sims:
        s.close()
    n = len(members)
    return {
        "workload": "64 infection sheets (sides by seed from 64..512) + 16 x 64^3 volumes, "
                    "J~U[5,15], lambda_V~U[1,4], T~U[5,20], seed 3",
        "lattices": n, "sites": sites, "mcs_each": mcs,
        "batched_replicas_per_s": n / batched, "looped_replicas_per_s": n / looped,
        "speedup": looped / batched, "batched_s": batched, "looped_s": looped,
        "batched_site_attempts_per_s": sites * mcs / batched,
        "batched_hbm_frac_algorithmic": algo / batched / (HBM_PEAK_GBS * 1e9),
        "batched_algorithmic_bytes": algo,
        "graph_step_ms": graph_step * 1e3, "lattices_identical": bool(same),
        "tiles": [{k: g[k] for k in ("dims", "lattices", "tile", "threads", "wgs_per_cu") if k in g}
                  for g in info["groups"]],
    }


def cpu_baseline_ensemble(steps: int = 10):
    """The serial oracle on the same 80-member ensemble, one member per host
    core at a time across every core: each core takes the next member and
    runs `steps` of its steps. Replicas/s extrapolates the wall time to the
    full 100 steps a replica runs."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    from glazier_amd.workloads import ensemble

    members = ensemble(seed=3)
    nproc = os.cpu_count() or 1
    # the cores this process may run on, at most 16 (one GPU's share of a
    # shared host)
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    sims = [oracle.SerialSimulation(m, l, t) for m, l, t in members]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as pool:
        list(pool.map(lambda sim: sim.step(steps), sims))  # ctypes drops the GIL in the call
    wall = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    return {"value": len(members) / (wall * 100.0 / steps), "unit": "replicas/s", "cores": cores,
            "nproc": nproc, "kind": "port",
            "sample": f"serial f64 oracle, all {len(members)} ensemble members for {steps} of their 100 "
                      f"steps on {cores} host threads (nproc {nproc}), one member per thread at a time, "
                      f"{wall:.2f} s, scaled to 100 steps"}


def adjoint_bench(cpu: bool, adam_steps: int = 50):
    """The differentiated field fit: a 256^2 infection sheet (1024 cells,
    seed 7) run for 200 steps with its owner trajectory recorded; targets
    from the replay with perturbed rates (D x1.3, decay x0.7, secretion x1.2,
    uptake x0.8); the gradient timed, checked against the f64 CPU adjoint,
    and used by Adam to bring the rates back."""
    import numpy as np
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.adjoint import FieldFit, fit_rates, record, recorded_types
    from glazier_amd.model import FieldParams
    from glazier_amd.workloads import infection_sheet

    steps, every = 200, 20
    model, labels, types = infection_sheet(256, 7)
    sim = HipSimulation(model, labels, types)
    record(sim, steps)
    sim.step(steps)
    fit = FieldFit(sim, checkpoint_every=every)
    base = fit.model_params()
    truth = FieldParams(base.diffusion * 1.3, base.decay * 0.7, base.secretion * 1.2, base.uptake * 0.8)
    u0 = torch.zeros((2, model.sites), dtype=torch.float32, device="cuda")
    target = fit.forward(truth, u0)
    # gradient timing: forward + backward of the L2 loss
    for _ in range(2):
        out = fit.forward(base, u0)
        fit.backward(2.0 * (out - target))
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fit.forward(base, u0)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        grad, _ = fit.backward(2.0 * (out - target))
    torch.cuda.synchronize()
    bwd = (time.perf_counter() - t0) / reps
    substeps = 4
    # algorithmic bytes of one replayed step: per species and substep one f32
    # read and one f32 write per site, plus the exchange's read, write and
    # one-byte owner read
    n = model.sites
    fwd_bytes = steps * 2 * n * (substeps * 8 + 8) + steps * n
    line = {
        "workload": "256^2 infection sheet, 1024 cells, seed 7, 200 steps recorded, "
                    "checkpoints every 20; rates perturbed D x1.3, decay x0.7, secretion x1.2, "
                    "uptake x0.8",
        "forward_ms": fwd * 1e3, "gradient_ms": (fwd + bwd) * 1e3, "backward_ms": bwd * 1e3,
        "forward_gbs_algorithmic": fwd_bytes / fwd / 1e9,
        "parameters": base.names(),
    }
    t0 = time.perf_counter()
    hist = fit_rates(fit, target, base, u0, steps=adam_steps, lr=0.2)
    torch.cuda.synchronize()
    free = base.vector() > 0
    rel = np.abs(hist[-1]["theta"][free] / truth.vector()[free] - 1.0)
    line["adam"] = {"steps": adam_steps, "lr": 0.2, "seconds": time.perf_counter() - t0,
                    "loss_start": hist[0]["loss"], "loss_end": hist[-1]["loss"],
                    "rate_rel_error_max": float(rel.max()), "rate_rel_error_mean": float(rel.mean()),
                    "start_rel_error_mean": float(np.abs(base.vector()[free] / truth.vector()[free]
                                                         - 1.0).mean())}
    # kept for the CPU leg: the trajectory and the device gradient
    line["_cpu_leg"] = {"model": model, "truth": truth, "base": base, "grad": grad,
                        "traj": np.stack([recorded_types(sim, k) for k in range(steps)]) if cpu else None}
    fit.close()
    sim.close()
    return line


def cpu_baseline_configs():
    """The serial oracle (one thread) on config 0 in full and on a bounded
    sample of config 2 (the first 10 of its 200 steps)."""
    import oracle
    from glazier_amd.workloads import infection_sheet, plain_monolayer

    out = {}
    for name, build, steps in (("config0_128", plain_monolayer(128, 1, False), 100),
                               ("config2_infection_512", infection_sheet(512, 7), 10)):
        sim = oracle.SerialSimulation(*build)
        t0 = time.perf_counter()
        sim.step(steps)
        el = time.perf_counter() - t0
        out[name] = {"value": build[0].sites * steps / el, "unit": "site-attempts/s", "cores": 1,
                     "kind": "port", "sample": f"{steps} steps in {el:.2f} s"}
    return out


def cpu_baseline_adjoint(leg: dict):
    """The f64 CPU adjoint (the oracle, one thread) of the same 256^2 fit:
    its time for one gradient, and how far the device gradient sits from it."""
    import numpy as np

    import oracle

    model, truth, base, grad, traj = leg["model"], leg["truth"], leg["base"], leg["grad"], leg["traj"]
    u0 = np.zeros((2, model.sites))
    target = oracle.replay_forward(model, truth, traj, u0)
    t0 = time.perf_counter()
    out = oracle.replay_forward(model, base, traj, u0)
    want, _ = oracle.replay_gradient(model, base, traj, u0, 2.0 * (out - target))
    seconds = time.perf_counter() - t0
    gv, wv = grad.vector(), want.vector()
    scale = np.abs(wv).max()
    return {"value": seconds, "unit": "s per gradient", "cores": 1, "kind": "port",
            "sample": "f64 replay and adjoint oracle, the whole 256^2 200-step fit, one thread",
            "device_vs_f64_max_rel": float(max(abs(gv[i] - wv[i]) / max(abs(wv[i]), 1e-3 * scale)
                                               for i in range(gv.size)))}


def configs_bench():
    """The other workloads, each from resident inputs: config 0 (128^2 plain,
    100 steps, seed 1), config 1 at 1024^2 and 2048^2 (100 steps, seeds
    1-4), config 2 (512^2 infection sheet, 200 steps, seed 7), and the
    reference's own benchmark model (examples/bench.rs:7-25: contact 16/8,
    V_t 64, lambda_V 2, T 10, Moore, seed 1, 100 steps) at the sizes whose
    device times the reference publishes."""
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.model import CellType, Model, default_cell_types, tiled_labels
    from glazier_amd.workloads import infection_sheet, plain_monolayer

    def run(build, steps, bytes_per_attempt):
        sims = [HipSimulation(*b) for b in build]
        for s in sims:
            s.step(1)
        torch.cuda.synchronize()
        acc0 = sum(s.counters()[1] for s in sims)
        t0 = time.perf_counter()
        for s in sims:
            s.step(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        sites = sims[0].sites
        attempts = sites * steps * len(sims)
        a = (sum(s.counters()[1] for s in sims) - acc0) / attempts
        for s in sims:
            s.close()
        return {"site_attempts_per_s": attempts / el, "mcs_per_s": steps * len(sims) / el,
                "acceptance": a, "seconds": el,
                "hbm_frac_algorithmic": attempts * bytes_per_attempt(a) / el / (HBM_PEAK_GBS * 1e9)}

    from glazier_amd.workloads import reference_model

    plain = lambda a: 16.0 + 4.0 * a
    # infection: 4 colours x (labels + the chemotactic field) + accepted
    # writes, then per site per step two fields read and written, the owner
    # read of the exchange, and the moments' label read
    infection = lambda a: 4.0 * (4.0 + 4.0) + 4.0 * a + 8.0 * 2 + 4.0 + 4.0
    out = {"config0_128": run([plain_monolayer(128, 1, False)], 100, plain)}
    for side in (1024, 2048):
        out[f"config1_{side}"] = run([plain_monolayer(side, seed, True) for seed in (1, 2, 3, 4)], 100, plain)
    out["config2_infection_512"] = run([infection_sheet(512, 7)], 200, infection)
    # README.md:116-121, site-attempts/s derived from the listed GPU seconds
    published = {128: 8.2e8, 256: 3.3e9, 512: 3.7e9, 1024: 4.4e9, 2048: 4.3e9, 4096: 3.76e9}
    for side, ref in published.items():
        r = run([reference_model(side, 1)], 100, plain)
        r["reference_published_site_attempts_per_s"] = ref
        r["ratio_to_reference"] = r["site_attempts_per_s"] / ref
        out[f"reference_model_{side}"] = r
    return out


def blueprints_bench():
    """The reference's published runs through the binary (README.md:206-210,
    248-250): the `glazier` command line on infection-512.json (200 steps),
    volume.json and infection-volume.json (64x64x32; 50 and 60 steps), each
    a child process from a cold start, as the reference times them: its
    `seconds` runs from before device setup (main.rs:153-157, 293, 317) and
    so does this binary's. Dumps are left as each blueprint asks."""
    import subprocess
    import tempfile

    root = Path(__file__).resolve().parent
    binary = root / "glazier_amd" / "glazier"
    published = {"infection-512.json": (9.72, 1.66), "volume.json": (1.60, 0.092),
                 "infection-volume.json": (1.09, 0.081)}
    out = {}
    for name, (cpu_s, dev_s) in published.items():
        with tempfile.TemporaryDirectory() as tmp:
            cmd = [str(binary), "--model", str(root / "tests" / "golden" / "blueprints" / name),
                   "--out", tmp, "--engine", "hip"]
            done = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            if done.returncode != 0:
                out[name] = {"error": done.stderr.strip()[-300:]}
                continue
            summary = json.loads((Path(tmp) / "summary.json").read_text())
        out[name] = {"steps": summary["steps"], "seconds": summary["seconds"],
                     "setup_seconds": summary["setup_seconds"],
                     "site_attempts_per_s": summary["site_attempts_per_second"],
                     "n_cells": summary["n_cells"],
                     "reference_device_seconds": dev_s, "reference_serial_seconds": cpu_s,
                     "speedup_vs_reference_device": dev_s / summary["seconds"]}
    return out


# README.md:153-162 and 170-178: the reference's glazier CPU and GPU columns
COMPARISON_PUBLISHED = {
    "monolayer-physicell.json": {
        "glazier_cpu": {"cells": 256, "mean_area": 64.0, "area_rms_from_target": 1.58,
                        "mean_eccentricity": 0.397, "nematic_order": 0.064, "defects": 56,
                        "net_charge": 0.0, "seconds": 0.146},
        "glazier_gpu": {"cells": 256, "mean_area": 64.0, "area_rms_from_target": 2.07,
                        "mean_eccentricity": 0.392, "nematic_order": 0.052, "defects": 50,
                        "net_charge": 0.0, "seconds": 0.106}},
    "monolayer-surface.json": {
        "glazier_cpu": {"mean_eccentricity": 0.371, "area_rms_from_target": 1.61, "defects": 34},
        "glazier_gpu": {"mean_eccentricity": 0.359, "area_rms_from_target": 1.72, "defects": 44}},
    "monolayer.json": {
        "glazier_cpu": {"mean_eccentricity": 0.394},
        "glazier_gpu": {"mean_eccentricity": 0.394}},
}


def comparison_bench():
    """The engine comparison (bench/engines.py, README.md:147-182) for this
    engine: each blueprint run for its 200 steps through the `glazier`
    binary, the final label field measured as the reference measures it
    (glazier_amd.director, coarse-grain sigma 4), the binary's own seconds."""
    import subprocess
    import tempfile

    import numpy as np

    from glazier_amd import director
    from glazier_amd.blueprint import Blueprint

    root = Path(__file__).resolve().parent
    binary = root / "glazier_amd" / "glazier"
    out = {}
    for name, published in COMPARISON_PUBLISHED.items():
        path = root / "tests" / "golden" / "blueprints" / name
        bp = Blueprint.load(path)
        with tempfile.TemporaryDirectory() as tmp:
            done = subprocess.run([str(binary), "--model", str(path), "--out", tmp, "--engine", "hip"],
                                  capture_output=True, text=True, timeout=300)
            if done.returncode != 0:
                out[name] = {"error": done.stderr.strip()[-300:]}
                continue
            summary = json.loads((Path(tmp) / "summary.json").read_text())
            labels = np.load(Path(tmp) / f"labels_{bp.steps:05d}.npy")
        row = director.measure(labels, bp.types[0].target_volume, 4.0)
        row["seconds"] = summary["seconds"]
        out[name] = {"hip": row, "reference_published": published}
    return out


def cpu_baseline_comparison():
    """The serial oracle's column of the engine comparison: the same
    blueprints and steps on one thread, measured the same way."""
    import numpy as np

    import oracle
    from glazier_amd import director
    from glazier_amd.blueprint import Blueprint

    root = Path(__file__).resolve().parent
    out = {}
    for name in COMPARISON_PUBLISHED:
        bp = Blueprint.load(root / "tests" / "golden" / "blueprints" / name)
        model, labels, types = bp.simulation()
        sim = oracle.SerialSimulation(model, labels, types)
        t0 = time.perf_counter()
        sim.step(bp.steps)
        el = time.perf_counter() - t0
        row = director.measure(sim.labels().reshape(model.shape()), bp.types[0].target_volume, 4.0)
        row["seconds"] = el
        out[name] = row
        sim.close()
    return out


def counter_per_dispatch(args, counter: str, kernel: str):
    """Run this script under `rocprofv3 --pmc <counter>` in a child process
    (the program directly after `--`) and return the counter's mean value per
    dispatch of `kernel`, summed over the counter's instances; None when the
    profiler is missing or the run fails."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    with tempfile.TemporaryDirectory() as out:
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "-d", out, "-o", "run", "--",
               sys.executable, str(ROOT / "bench.py"), *args]
        done = subprocess.run(cmd, capture_output=True, text=True)
        dbs = sorted(Path(out).rglob("*.db"))
        if not dbs:
            return None
        con = sqlite3.connect(str(dbs[0]))
        rows = con.execute("select dispatch_id, sum(counter_value) from pmc_events where name like ? and "
                           "counter_name = ? group by dispatch_id", (f"%{kernel}%", counter)).fetchall()
        con.close()
        del done
    if not rows:
        return None
    return sum(r[1] for r in rows) / len(rows)


def measured_traffic(side: int):
    """HBM bytes per launch of the sweep pass from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by what the same counter reads for a
    256 MiB torch copy (256 MiB read, 256 MiB written)."""
    probe = ["--traffic-probe", "--side", str(side)]
    copy = ["--copy-probe", "256"]
    f_sweep = counter_per_dispatch(probe, "FETCH_SIZE", "sweep_plain2d")
    w_sweep = counter_per_dispatch(probe, "WRITE_SIZE", "sweep_plain2d")
    f_copy = counter_per_dispatch(copy, "FETCH_SIZE", "copyBuffer")
    w_copy = counter_per_dispatch(copy, "WRITE_SIZE", "copyBuffer")
    if None in (f_sweep, w_sweep, f_copy, w_copy) or not f_copy or not w_copy:
        return None
    true_kib = 256 * 1024
    f_scale, w_scale = true_kib / f_copy, true_kib / w_copy
    return {"bytes": (f_sweep * f_scale + w_sweep * w_scale) * 1024.0,
            "read_bytes": f_sweep * f_scale * 1024.0, "write_bytes": w_sweep * w_scale * 1024.0,
            "fetch_scale": f_scale, "write_scale": w_scale}


def moving_roofline(side: int):
    """The same pass on the reference's benchmark model (examples/bench.rs:
    7-25: contact 16/8, no surface term), whose sheet moves (about 7% of
    attempts accepted): one event pair around 25 steps (100 back-to-back
    passes) on the engine's stream, bytes per launch 4 N + 4 per accepted
    copy."""
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.workloads import reference_model

    sim = HipSimulation(*reference_model(side, 1))
    sim.step(8)
    torch.cuda.synchronize()
    acc0 = sim.counters()[1]
    sim.set_timing(2)
    sim.step(25)
    sim.set_timing(False)
    ms, launches = sim.kernel_time("step")
    acc = sim.counters()[1] - acc0
    sim.close()
    n = side * side
    avg_s = ms / launches / 1e3
    bytes_per_launch = 4.0 * n + 4.0 * acc / launches
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "kernel": "sweep_plain2d_kernel (one colour pass)",
            "workload": f"reference benchmark model {side}^2, 8x8 cells, contact 16/8, V_t=64, "
                        "lambda_V=2, T=10, Moore, seed 1",
            "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_per_launch": bytes_per_launch,
            "acceptance": acc / (25.0 * n),
            "avg_launch_us_method": "HIP events around 100 back-to-back launches on the engine's "
                                    "stream, divided by 100"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-mcs", type=int, default=100)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure the pass's HBM bytes")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run a few passes of one lattice (the counter runs use this)")
    ap.add_argument("--probe", choices=["config2", "moving4096", "ref256", "ref512"], default=None,
                    help="only step one workload and print its step time, for a profiler run: "
                         "config2 (512^2 infection sheet, seed 7, 200 steps), moving4096 "
                         "(the reference's benchmark model at 4096^2, 25 steps), ref256 or ref512 "
                         "(that model at 256^2 or 512^2, 100 steps)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-mebibyte buffer 5 times with torch, so a counter "
                         "run can scale FETCH_SIZE/WRITE_SIZE by a copy of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build",
              file=sys.stderr)
        return 2

    import torch  # device memory plumbing only: the engine allocates its own

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "copies": 5}))
        return 0
    from glazier_amd import HipSimulation
    from glazier_amd.workloads import plain_monolayer

    side = args.side
    if args.probe:
        from glazier_amd.workloads import infection_sheet, reference_model

        built, steps = {"config2": lambda: (infection_sheet(512, 7), 200),
                        "moving4096": lambda: (reference_model(4096, 1), 25),
                        "ref256": lambda: (reference_model(256, 1), 100),
                        "ref512": lambda: (reference_model(512, 1), 100)}[args.probe]()
        sim = HipSimulation(*built)
        sim.step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        # the same steps again with HIP events around every sweep launch
        sim.set_timing(True)
        sim.step(steps)
        sim.set_timing(False)
        ev_ms, ev_launches = sim.kernel_time("sweep")
        print(json.dumps({"probe": args.probe, "steps": steps, "us_per_step": el / steps * 1e6,
                          "n_labels": sim.n_labels(), "sweep_launches": ev_launches,
                          "sweep_event_us_per_step": ev_ms * 1e3 / steps}))
        sim.close()
        return 0
    if args.traffic_probe:
        model, labels, types = plain_monolayer(side, seed=1, surface=True)
        sim = HipSimulation(model, labels, types)
        sim.step(4)
        torch.cuda.synchronize()
        sim.close()
        return 0
    sims = []
    for seed in (1, 2, 3, 4):
        model, labels, types = plain_monolayer(side, seed=seed, surface=True)
        sims.append(HipSimulation(model, labels, types))
    for i in range(args.warmup):
        sims[i % 4].step(1)
    torch.cuda.synchronize()
    acc0 = [s.counters()[1] for s in sims]

    per = [args.steps // 4 + (1 if i < args.steps % 4 else 0) for i in range(4)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, k in zip(sims, per):
        if k:
            s.step(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    n = side * side
    attempts = n * args.steps
    accepted = sum(s.counters()[1] for s in sims) - sum(acc0)
    a = accepted / attempts

    # roofline of the sweep kernel, timed with HIP events on the engine's own
    # stream over a separate stretch of the same run: one event pair around
    # a call of 25 steps (100 back-to-back sweep launches; a step of this
    # workload is its four colour passes and nothing else), and, for
    # comparison, an event pair around every launch (which adds the events'
    # own cost between launches)
    probe = sims[0]
    before_acc = probe.counters()[1]
    probe.set_timing(2)
    probe.step(25)
    probe.set_timing(False)
    ms, launches = probe.kernel_time("step")
    acc_probe = probe.counters()[1] - before_acc
    probe.set_timing(True)
    probe.step(10)
    probe.set_timing(False)
    ms_b, launches_b = probe.kernel_time("sweep")
    avg_s = ms / launches / 1e3
    bytes_per_launch = 4.0 * n + 4.0 * acc_probe / launches
    achieved = bytes_per_launch / avg_s / 1e9

    line = {
        "metric": "site-attempts/s at 4096^2 (% HBM peak); batched vs looped replicas/s",
        "value": attempts / elapsed,
        "unit": "site-attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 labels, int64 counters, f32 energies",
        "data": "synthetic",
        "config": {"workload": f"plain monolayer {side}^2, 8x8 cells, J=10, lambda_V=2, V_t=64, "
                               f"lambda_S=1, S_t=32, T=10, Moore; seeds 1-4 round robin",
                   "side": side, "cells": (side // 8) ** 2, "seeds": [1, 2, 3, 4],
                   "mcs_per_seed": per},
        "mcs_per_s": args.steps / elapsed,
        "acceptance": a,
        "hbm_frac_algorithmic": attempts * (16.0 + 4.0 * a) / elapsed / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "sweep_plain2d_kernel (one colour pass)", "traffic_unit": "bytes per launch",
                     "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_us_method": "HIP events around 100 back-to-back launches on the engine's "
                                             "stream, divided by 100",
                     "avg_launch_us_each_bracketed": ms_b / launches_b * 1e3},
        "cpu_baseline": None,
    }
    for s in sims:
        s.close()
    line["roofline_moving"] = moving_roofline(side)
    if not args.no_traffic:
        traffic = measured_traffic(side)
        if traffic is not None:
            line["roofline"]["traffic"] = traffic["bytes"]
            line["roofline"]["traffic_detail"] = traffic
    if not args.no_configs:
        line["configs"] = configs_bench()
        line["blueprints"] = blueprints_bench()
        line["engine_comparison"] = comparison_bench()
    if not args.no_ensemble:
        line["ensemble"] = ensemble_bench(args.ensemble_mcs)
    leg = None
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench(cpu=not args.no_cpu)
        leg = line["adjoint"].pop("_cpu_leg")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(side, args.cpu_seconds)
        if leg is not None:
            line["cpu_baseline"]["adjoint"] = cpu_baseline_adjoint(leg)
        if not args.no_ensemble:
            line["cpu_baseline"]["ensemble"] = cpu_baseline_ensemble()
        if not args.no_configs:
            line["cpu_baseline"]["configs"] = cpu_baseline_configs()
            line["cpu_baseline"]["engine_comparison"] = cpu_baseline_comparison()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

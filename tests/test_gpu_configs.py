"""Every measured workload at its stated size, on the device, against the
CPU emulation of the same schedule (bit for bit) or the f64 references.

* Plain sweep at scale: 2048^2 and 4096^2 periodic sheets of 8 x 8 cells,
  J = 10, lambda_V = 2, V_t = 64, lambda_S = 1, S_t = 32, T = 10, Moore,
  seeds 1-4 (the headline workload; 3 and 2 steps per seed).
* The infection sheet: 512^2, 4096 cells, two types, two f32 fields with
  four substeps, chemotaxis, length, death and division, seed 7 (25 steps),
  and the same physics with division and death turned up so the run crosses
  divisions, deaths and two compactions.
* The ensemble: 64 infection sheets of sides 64-512 and 16 volumes of 64^3
  (26 neighbours), stepped in one batch through its nine tuning steps and
  two graph replays, every member against its own emulation.
* The field fit: a 256^2 infection sheet (1024 cells) recorded over 200
  steps, checkpoints every 20: the device gradient against the f64 adjoint
  and f64 central differences (h = 1e-4 relative) on four rates.
"""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from glazier_amd.model import FieldParams

pytestmark = pytest.mark.gpu


def same(gpu, emu, what):
    assert gpu.n_labels() == emu.n_labels(), what
    diff = int((gpu.labels() != emu.labels()).sum())
    assert diff == 0, f"{what}: {diff} sites differ"
    assert np.array_equal(gpu.cell_state()[1:], emu.cell_state()[1:]), what
    assert np.array_equal(gpu.cell_types(), emu.cell_types()), what
    assert gpu.counters() == emu.counters(), what
    for k in range(len(gpu.model.species)):
        assert np.array_equal(gpu.field_f32(k), emu.field_f32(k)), f"{what}: species {k}"


def run_pair(model, labels, types, steps, chunks=None):
    from glazier_amd import HipSimulation

    gpu = HipSimulation(model, labels, types)
    with ThreadPoolExecutor(max_workers=1) as pool:
        def emulate():
            emu = oracle.EmuSimulation(model, labels, types)
            emu.step(steps)
            return emu

        pending = pool.submit(emulate)
        for c in chunks or [steps]:
            gpu.step(c)
        emu = pending.result()
    return gpu, emu


@pytest.mark.parametrize("side,steps", [(2048, 3), (4096, 2)])
def test_the_plain_sweep_at_scale_matches_the_emulation(engine, oracle_lib, side, steps):
    from glazier_amd.workloads import plain_monolayer

    for seed in (1, 2, 3, 4):
        print(f"[config] plain {side}^2 seed {seed}, {steps} steps", flush=True)
        m, labels, types = plain_monolayer(side, seed, True)
        gpu, emu = run_pair(m, labels, types, steps, chunks=[1, steps - 1])
        same(gpu, emu, f"{side}^2 seed {seed}")
        gpu.close()


def test_the_infection_sheet_matches_the_emulation(engine, oracle_lib):
    from glazier_amd.workloads import infection_sheet

    m, labels, types = infection_sheet(512, 7)
    print("[config] infection 512^2 seed 7, 25 steps", flush=True)
    gpu, emu = run_pair(m, labels, types, 25, chunks=[5, 20])
    same(gpu, emu, "infection 512")
    assert (emu.volumes()[1:] == 0).any(), "no cell died"
    gpu.close()


def test_the_infection_physics_through_divisions_and_compactions(engine, oracle_lib):
    from glazier_amd.workloads import infection_sheet

    m, labels, types = infection_sheet(512, 7)
    m.types[2].division_volume = 66.0
    m.types[1].death_rate = 0.04
    print("[config] infection 512^2 with events, 25 steps", flush=True)
    gpu, emu = run_pair(m, labels, types, 25, chunks=[7, 1, 17])
    same(gpu, emu, "infection 512 with events")
    # 4097 labels at the start; compaction brought the count below it
    assert emu.n_labels() < 3000
    gpu.close()


def test_the_full_ensemble_matches_each_members_emulation(engine, oracle_lib):
    from glazier_amd import HipBatch, HipSimulation
    from glazier_amd.workloads import ensemble

    members = ensemble(seed=3)
    assert len(members) == 80 and sum(m.depth == 64 for m, _, _ in members) == 16
    sims = [HipSimulation(m, l, t) for m, l, t in members]
    steps = 11
    print(f"[config] ensemble of {len(sims)}, {steps} steps", flush=True)

    def emulate(member):
        m, l, t = member
        emu = oracle.EmuSimulation(m, l, t)
        emu.step(steps)
        return emu

    with ThreadPoolExecutor(max_workers=8) as pool:
        pending = pool.map(emulate, members)
        batch = HipBatch(sims)
        batch.step(steps)
        info = batch.describe()
        emus = list(pending)
    assert info["tuned"] and info["eager_steps"] == 9 and info["graph_steps"] == 2
    for (m, _, _), gpu, emu in zip(members, sims, emus):
        same(gpu, emu, f"replica {m.replica} ({m.width}x{m.height}x{m.depth})")
    batch.close()
    for s in sims:
        s.close()


def test_the_field_fit_gradient_at_its_stated_size(engine, oracle_lib):
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.adjoint import FieldFit, record, recorded_types
    from glazier_amd.workloads import infection_sheet

    steps, every = 200, 20
    m, labels, types = infection_sheet(256, 7)
    sim = HipSimulation(m, labels, types)
    record(sim, steps)
    print("[config] field fit 256^2, 200 steps", flush=True)
    sim.step(steps)
    traj = np.stack([recorded_types(sim, s) for s in range(steps)])
    fit = FieldFit(sim, checkpoint_every=every)
    base = fit.model_params()
    truth = FieldParams(base.diffusion * 1.3, base.decay * 0.7, base.secretion * 1.2, base.uptake * 0.8)
    rng = np.random.default_rng(11)
    u0 = rng.uniform(0.1, 1.0, size=(len(m.species), m.sites))
    target = oracle.replay_forward(m, truth, traj, u0)

    out = fit.forward(base, torch.tensor(u0, dtype=torch.float32, device="cuda"))
    grad, _ = fit.backward(2.0 * (out - torch.tensor(target, dtype=torch.float32, device="cuda")))
    got = grad.vector()

    out64 = oracle.replay_forward(m, base, traj, u0)
    want, _ = oracle.replay_gradient(m, base, traj, u0, 2.0 * (out64 - target))
    want = want.vector()
    scale = np.abs(want).max()
    names = base.names()
    for i, nm in enumerate(names):
        assert abs(got[i] - want[i]) <= 1e-3 * max(abs(want[i]), 1e-3 * scale), (nm, got[i], want[i])

    def loss(v):
        p = FieldParams.from_vector(v, base.n_types, len(m.species))
        return float(((oracle.replay_forward(m, p, traj, u0) - target) ** 2).sum())

    v = base.vector()
    free = [i for i in range(len(v)) if v[i] != 0.0]
    # a diffusion, a decay, a secretion and an uptake rate
    chosen = []
    for kind in ("diffusion", "decay", "secretion", "uptake"):
        chosen.append(next(i for i in free if names[i].startswith(kind)))
    for i in chosen:
        h = 1e-4 * abs(v[i])
        up, dn = v.copy(), v.copy()
        up[i] += h
        dn[i] -= h
        fd = (loss(up) - loss(dn)) / (2 * h)
        print(f"[config] d loss / d {names[i]}: device {got[i]:.6e}, f64 differences {fd:.6e}", flush=True)
        assert abs(got[i] - fd) <= 1e-3 * max(abs(fd), 1e-3 * scale), (names[i], got[i], fd)
    sim.close()

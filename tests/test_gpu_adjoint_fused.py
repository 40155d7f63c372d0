"""The one-launch-per-step replay and adjoint on 2D lattices against the
per-substep launches (GLAZIER_FIT_PER_SUBSTEP=1).

Both run the same f32 operations per site in the same order, so the replayed
fields and the adjoint of the initial fields must agree bit for bit; the
parameter gradients hold the same f64 terms summed in another grouping and
must agree to 1e-12 relative. The lattices tile unevenly (70 x 46 against
32 x 16 tiles) or are smaller than one tile and its halo (20 x 12), and the
two species take different substep counts (4 and 2) so one idles through
part of the step.
"""

import numpy as np
import pytest

from glazier_amd.model import FieldParams, Species
from tests.cases import ref_model, tiled

pytestmark = pytest.mark.gpu


def fit_pair(monkeypatch, w, h, steps, every):
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.adjoint import FieldFit, record

    m = ref_model(w, h, seed=17)
    m.species = [Species("a", 0.8, 0.02, 0.0, substeps=4), Species("b", 0.3, 0.05, 0.0, substeps=2)]
    m.secretion = [[0.0, 0.0], [0.4, 0.15]]
    m.uptake = [[0.0, 0.0], [0.05, 0.3]]
    m.chemotaxis = [[0.0, 0.0], [0.0, 0.0]]
    labels, types = tiled(m, side=6, nx=max(1, w // 8), ny=max(1, h // 8))
    sim = HipSimulation(m, labels, types)
    record(sim, steps)
    sim.step(steps)
    monkeypatch.delenv("GLAZIER_FIT_PER_SUBSTEP", raising=False)
    fused = FieldFit(sim, checkpoint_every=every)
    monkeypatch.setenv("GLAZIER_FIT_PER_SUBSTEP", "1")
    per_substep = FieldFit(sim, checkpoint_every=every)
    monkeypatch.delenv("GLAZIER_FIT_PER_SUBSTEP")
    return m, sim, fused, per_substep, torch


@pytest.mark.parametrize("w,h,steps,every", [(70, 46, 14, 5), (20, 12, 9, 4), (128, 64, 6, 6)])
def test_one_launch_per_step_matches_the_per_substep_launches(monkeypatch, engine, w, h, steps, every):
    print(f"[fused adjoint] {w}x{h}, {steps} steps, checkpoints every {every}", flush=True)
    m, sim, fused, per_substep, torch = fit_pair(monkeypatch, w, h, steps, every)
    base = fused.model_params()
    zero = torch.zeros((2, m.sites), dtype=torch.float32, device="cuda")
    # the engine's own rates from zero fields: the engine's fields, bit for bit
    out = fused.forward(base, zero).cpu().numpy()
    for k in range(2):
        assert np.array_equal(out[k], sim.field_f32(k)), f"species {k}"

    theta = FieldParams(base.diffusion * 1.1, base.decay * 0.9, base.secretion * 1.2, base.uptake * 0.7)
    rng = np.random.default_rng(5)
    u0 = torch.tensor(rng.uniform(0.0, 2.0, size=(2, m.sites)), dtype=torch.float32, device="cuda")
    target = torch.tensor(rng.uniform(0.0, 2.0, size=(2, m.sites)), dtype=torch.float32, device="cuda")
    a = fused.forward(theta, u0)
    b = per_substep.forward(theta, u0)
    assert torch.equal(a, b)
    ga, ga_u0 = fused.backward(2.0 * (a - target))
    gb, gb_u0 = per_substep.backward(2.0 * (b - target))
    assert torch.equal(ga_u0, gb_u0)
    va, vb = ga.vector(), gb.vector()
    scale = np.abs(vb).max()
    assert scale > 0
    for i, nm in enumerate(base.names()):
        assert abs(va[i] - vb[i]) <= 1e-12 * max(abs(vb[i]), scale), (nm, va[i], vb[i])

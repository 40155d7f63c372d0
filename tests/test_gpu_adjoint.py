"""The device field replay and its adjoint.

The replay of a lattice's own rates over its recorded owners must reproduce
the engine's fields bit for bit; the device gradient must agree with the f64
CPU adjoint and with f64 central differences of the f64 replay (h = 1e-4
relative) within 1e-3; the torch Function must hand the same gradients to
autograd; and Adam on the gradient must bring perturbed rates back.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import FieldParams
from tests.cases import case

pytestmark = pytest.mark.gpu


def recorded(name, steps):
    from glazier_amd import HipSimulation
    from glazier_amd.adjoint import record

    m, labels, types, fields, activity, _ = case(name)
    sim = HipSimulation(m, labels, types, fields, activity)
    record(sim, steps)
    sim.step(steps)
    return m, labels, types, fields, activity, sim


def trajectory(sim, steps):
    from glazier_amd.adjoint import recorded_types

    return np.stack([recorded_types(sim, s) for s in range(steps)])


def test_the_recording_holds_each_steps_owners(engine, oracle_lib):
    m, labels, types, fields, activity, sim = recorded("secreting", 12)
    emu = oracle.EmuSimulation(m, labels, types, fields, activity)
    traj = trajectory(sim, 12)
    for s in range(12):
        emu.step(1)
        want = emu.cell_types()[emu.labels()]
        want[emu.labels() == 0] = 0
        assert np.array_equal(traj[s], want), f"step {s}"


@pytest.mark.parametrize("every", [7, 30])
def test_the_replay_reproduces_the_engine_fields(engine, oracle_lib, every):
    import torch

    from glazier_amd.adjoint import FieldFit

    m, labels, types, fields, activity, sim = recorded("infection128", 30)
    fit = FieldFit(sim, checkpoint_every=every)
    u0 = torch.zeros((2, m.sites), dtype=torch.float32, device="cuda")
    out = fit.forward(fit.model_params(), u0).cpu().numpy()
    for k in range(2):
        assert np.array_equal(out[k], sim.field_f32(k)), f"species {k}"


def perturbed(base):
    return FieldParams(base.diffusion * 1.3, base.decay * 0.7, base.secretion * 1.2, base.uptake * 0.8)


@pytest.mark.parametrize("name,steps,every", [("infection128", 25, 7), ("division3d", 6, 4)])
def test_the_device_gradient_matches_the_f64_adjoint_and_differences(engine, oracle_lib, name, steps,
                                                                     every):
    import torch

    from glazier_amd.adjoint import FieldFit

    if name == "division3d":
        m, labels, types, fields, activity, _ = case(name)
        from glazier_amd.model import Species

        m.species = [Species("a", 0.1, 0.01, 0.0, substeps=3), Species("b", 0.05, 0.02, 0.0, substeps=2)]
        m.secretion = [[0.0, 0.0], [0.5, 0.2]]
        m.uptake = [[0.0, 0.0], [0.05, 0.1]]
        m.chemotaxis = [[0.0, 0.0], [0.0, 0.0]]
        from glazier_amd import HipSimulation
        from glazier_amd.adjoint import record

        sim = HipSimulation(m, labels, types)
        record(sim, steps)
        sim.step(steps)
    else:
        m, labels, types, fields, activity, sim = recorded(name, steps)
    traj = trajectory(sim, steps)
    fit = FieldFit(sim, checkpoint_every=every)
    base = fit.model_params()
    rng = np.random.default_rng(3)
    u0 = rng.uniform(0.1, 1.0, size=(len(m.species), m.sites))
    target = oracle.replay_forward(m, perturbed(base), traj, u0)

    out = fit.forward(base, torch.tensor(u0, dtype=torch.float32, device="cuda"))
    g_out = 2.0 * (out - torch.tensor(target, dtype=torch.float32, device="cuda"))
    grad, grad_u0 = fit.backward(g_out)

    out64 = oracle.replay_forward(m, base, traj, u0)
    want, want_u0 = oracle.replay_gradient(m, base, traj, u0, 2.0 * (out64 - target))
    got_v, want_v = grad.vector(), want.vector()
    scale = np.abs(want_v).max()
    for i, nm in enumerate(base.names()):
        assert abs(got_v[i] - want_v[i]) <= 1e-3 * max(abs(want_v[i]), 1e-3 * scale), (nm, got_v[i], want_v[i])
    gu = grad_u0.cpu().numpy()
    assert np.abs(gu - want_u0).max() <= 1e-3 * np.abs(want_u0).max()

    def loss(p):
        return float(((oracle.replay_forward(m, p, traj, u0) - target) ** 2).sum())

    v = base.vector()
    n_t = base.n_types
    for i, nm in enumerate(base.names()):
        if v[i] == 0.0:
            continue
        h = 1e-4 * abs(v[i])
        up, dn = v.copy(), v.copy()
        up[i] += h
        dn[i] -= h
        fd = (loss(FieldParams.from_vector(up, n_t, len(m.species)))
              - loss(FieldParams.from_vector(dn, n_t, len(m.species)))) / (2 * h)
        assert abs(got_v[i] - fd) <= 1e-3 * max(abs(fd), 1e-3 * scale), (nm, got_v[i], fd)


def test_the_torch_function_hands_autograd_the_adjoint(engine, oracle_lib):
    import torch

    from glazier_amd.adjoint import FieldFit, replay

    m, labels, types, fields, activity, sim = recorded("infection128", 10)
    fit = FieldFit(sim, checkpoint_every=4)
    base = fit.model_params()
    u0 = torch.rand((2, m.sites), dtype=torch.float32, device="cuda")
    target = torch.rand((2, m.sites), dtype=torch.float32, device="cuda")
    theta = torch.tensor(base.vector(), dtype=torch.float64, requires_grad=True)
    u0g = u0.clone().requires_grad_(True)
    out = replay(fit, theta, u0g)
    # another forward pass in between: the backward reruns its own
    fit.forward(perturbed(base), u0)
    ((out - target) ** 2).sum().backward()
    fit.forward(base, u0)
    want, want_u0 = fit.backward(2.0 * (out.detach() - target))
    assert np.allclose(theta.grad.numpy(), want.vector(), rtol=1e-9, atol=0.0)
    assert torch.equal(u0g.grad, want_u0)


def test_adam_brings_perturbed_rates_back(engine, oracle_lib):
    import torch

    from glazier_amd.adjoint import FieldFit, fit_rates

    m, labels, types, fields, activity, sim = recorded("infection128", 40)
    fit = FieldFit(sim, checkpoint_every=20)
    base = fit.model_params()
    truth = perturbed(base)
    u0 = torch.zeros((2, m.sites), dtype=torch.float32, device="cuda")
    target = fit.forward(truth, u0)
    hist = fit_rates(fit, target, base, u0, steps=50, lr=0.2)
    print(f"[adam] loss {hist[0]['loss']:.4e} -> {hist[-1]['loss']:.4e}", flush=True)
    assert hist[-1]["loss"] < 1e-2 * hist[0]["loss"]
    free = base.vector() > 0
    rel = np.abs(hist[-1]["theta"][free] / truth.vector()[free] - 1.0)
    start = np.abs(base.vector()[free] / truth.vector()[free] - 1.0)
    assert rel.mean() < 0.5 * start.mean()

"""The f64 field replay and its adjoint, pinned on the CPU.

The replay must be the serial engine's own field step (glazier-core's
Fields::diffuse and Fields::exchange as oracle/serial.c restates them) run
over a given sequence of owner maps; its adjoint must agree with central
finite differences of that replay, for every parameter and for the initial
field, in planes and volumes.
"""

import numpy as np
import pytest

import oracle
from glazier_amd.model import CellType, FieldParams, Model, Species


def small_model(w=16, h=16, d=1, substeps=(3, 2)):
    types = [CellType(), CellType(target_volume=16.0, lambda_volume=1.0),
             CellType(target_volume=16.0, lambda_volume=1.0)]
    return Model(width=w, height=h, depth=d, contact=[0.0] * 9, types=types, temperature=10.0,
                 neighbour_order=1, seed=3,
                 species=[Species("a", 0.12, 0.03, 0.0, substeps=substeps[0]),
                          Species("b", 0.07, 0.05, 0.0, substeps=substeps[1])],
                 secretion=[[0.0, 0.0], [0.3, 0.0], [0.8, 0.45]],
                 uptake=[[0.0, 0.0], [0.05, 0.02], [0.0, 0.1]],
                 chemotaxis=[[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]])


def trajectory(model, steps, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 3, size=(steps, model.sites), dtype=np.uint8)


def test_the_replay_is_the_serial_field_step(oracle_lib):
    m = small_model()
    types = trajectory(m, 6, 1)
    rng = np.random.default_rng(2)
    u0 = rng.uniform(0.0, 2.0, size=(2, m.sites))
    # labels that are their own type: the serial step then reads the map
    serial = oracle.SerialSimulation(m, types[0].astype(np.uint32),
                                     np.array([0, 1, 2], dtype=np.uint8), fields=u0)
    for step in range(types.shape[0]):
        serial.set_labels(types[step].astype(np.uint32), recount=False)
        serial.fields_step()
    want = np.stack([serial.field(k) for k in range(2)])
    got = oracle.replay_forward(m, FieldParams.from_model(m), types, u0)
    assert np.array_equal(got, want)


def loss_and_grad(m, params, types, u0, target):
    out = oracle.replay_forward(m, params, types, u0)
    return float(((out - target) ** 2).sum()), 2.0 * (out - target)


@pytest.mark.parametrize("dims", [(16, 16, 1), (6, 8, 4)])
def test_the_adjoint_matches_central_differences(oracle_lib, dims):
    m = small_model(*dims)
    types = trajectory(m, 8, 5)
    rng = np.random.default_rng(6)
    u0 = rng.uniform(0.2, 1.5, size=(2, m.sites))
    base = FieldParams.from_model(m)
    truth = FieldParams.from_vector(base.vector() * rng.uniform(0.7, 1.3, base.vector().size),
                                    3, 2)
    target = oracle.replay_forward(m, truth, types, u0)
    _, g_out = loss_and_grad(m, base, types, u0, target)
    grad, grad_u0 = oracle.replay_gradient(m, base, types, u0, g_out)
    v, gv = base.vector(), grad.vector()
    for i in range(v.size):
        h = 1e-6 * max(abs(v[i]), 1e-3)
        up, dn = v.copy(), v.copy()
        up[i] += h
        dn[i] -= h
        fd = (loss_and_grad(m, FieldParams.from_vector(up, 3, 2), types, u0, target)[0]
              - loss_and_grad(m, FieldParams.from_vector(dn, 3, 2), types, u0, target)[0]) / (2 * h)
        assert abs(gv[i] - fd) <= 1e-6 * max(abs(fd), 1e-3), (base.names()[i], gv[i], fd)
    for k, site in [(0, 3), (1, m.sites // 2), (1, m.sites - 1)]:
        h = 1e-6
        up, dn = u0.copy(), u0.copy()
        up[k, site] += h
        dn[k, site] -= h
        fd = (loss_and_grad(m, base, types, up, target)[0]
              - loss_and_grad(m, base, types, dn, target)[0]) / (2 * h)
        assert abs(grad_u0[k, site] - fd) <= 1e-6 * max(abs(fd), 1e-3)


def test_a_dropped_term_is_caught(oracle_lib):
    """The same check against a gradient with the decay term removed from
    the transposed substep: it must fail."""
    m = small_model()
    types = trajectory(m, 5, 7)
    u0 = np.random.default_rng(8).uniform(0.2, 1.5, size=(2, m.sites))
    base = FieldParams.from_model(m)
    grad, _ = oracle.replay_gradient(m, base, types, u0, np.ones((2, m.sites)))
    no_decay = FieldParams.from_model(m)
    no_decay.decay[:] = 0.0
    # the derivative with respect to D of sum(u_final) at zero decay differs
    grad_wrong, _ = oracle.replay_gradient(m, no_decay, types, u0, np.ones((2, m.sites)))
    h = 1e-6
    up, dn = base.vector(), base.vector()
    up[0] += h
    dn[0] -= h
    fd = (oracle.replay_forward(m, FieldParams.from_vector(up, 3, 2), types, u0).sum()
          - oracle.replay_forward(m, FieldParams.from_vector(dn, 3, 2), types, u0).sum()) / (2 * h)
    assert abs(grad.diffusion[0] - fd) <= 1e-6 * abs(fd)
    assert abs(grad_wrong.diffusion[0] - fd) > 1e-4 * abs(fd)

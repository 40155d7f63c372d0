"""The backward pass that allocates nothing (ghip_fit_backward_noalloc)
against ghip_fit_backward, through the C ABI.

The variant keeps its partial sums in the library's static device scratch
instead of the fit's own buffers and clears them itself: two calls in a row,
the second on a smaller lattice (fewer partial slots than the first left
behind), must each give what the allocating entry point gives, bit for bit.
"""

import ctypes

import numpy as np
import pytest

from glazier_amd.model import CFieldParams, FieldParams

pytestmark = pytest.mark.gpu


def fit_for(side, steps, every, seed):
    import torch

    from glazier_amd import HipSimulation
    from glazier_amd.adjoint import FieldFit, record
    from glazier_amd.workloads import infection_sheet

    m, labels, types = infection_sheet(side, seed)
    sim = HipSimulation(m, labels, types)
    record(sim, steps)
    sim.step(steps)
    fit = FieldFit(sim, checkpoint_every=every)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    u0 = torch.rand((2, m.sites), dtype=torch.float32, device="cuda", generator=gen)
    target = torch.rand((2, m.sites), dtype=torch.float32, device="cuda", generator=gen)
    out = fit.forward(fit.model_params(), u0)
    return sim, fit, 2.0 * (out - target)


def test_the_static_scratch_backward_matches_the_allocating_one(engine):
    import torch

    from glazier_amd.adjoint import _bind, _stream

    lib = _bind()
    for side, steps, every, seed in ((256, 24, 10, 7), (64, 9, 4, 3)):
        print(f"[no-alloc] {side}^2, {steps} steps, checkpoints every {every}", flush=True)
        sim, fit, g = fit_for(side, steps, every, seed)
        want, want_u0 = fit.backward(g)
        gc = CFieldParams()
        gu0 = torch.empty_like(g)
        status = lib.ghip_fit_backward_noalloc(fit._h, ctypes.c_void_p(g.data_ptr()), ctypes.byref(gc),
                                               ctypes.c_void_p(gu0.data_ptr()), _stream())
        assert status == 0, lib.ghip_last_error().decode()
        got = FieldParams.from_c(gc, fit.n_types, fit.n_species)
        assert np.array_equal(got.vector(), want.vector()), (got.vector(), want.vector())
        assert torch.equal(gu0, want_u0)
        fit.close()
        sim.close()

"""Differentiable fixed-weight solve (WLSSolve) on the GPU against the FP64
oracle (forward and adjoint), torch.autograd.gradcheck in FP64 and
central finite differences through the learnable coherence-to-weight
mapping."""

import numpy as np
import pytest

import oracle as O
from cuphu_amd import synth

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def _torch():
    import torch

    return torch


def _case(nrow, ncol, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.05, 0.5, O.narcs(nrow, ncol))
    phi = rng.uniform(0, TWOPI, (nrow, ncol)).astype(np.float32)
    return w, phi, rng


@pytest.mark.parametrize("shape", [(6, 7), (37, 53), (128, 128)])
def test_wls_forward_and_adjoint_match_fp64_oracle(gpu_ready, shape):
    torch = _torch()
    from cuphu_amd.autograd import WLSSolve

    w, phi, rng = _case(*shape, seed=shape[0])
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    pt = torch.tensor(phi, device="cuda")
    u = WLSSolve.apply(wt, pt, 20000, 1e-13)
    uo, _ = O.wls_solve_f64(w, phi.astype(np.float64))
    np.testing.assert_allclose(u.detach().cpu().numpy(), uo, atol=1e-8)
    gu = rng.normal(0, 1, shape)
    u.backward(torch.tensor(gu, device="cuda"))
    gwo, _ = O.wls_adjoint_f64(w, phi.astype(np.float64), uo, gu)
    np.testing.assert_allclose(wt.grad.cpu().numpy(), gwo, rtol=1e-6, atol=1e-8 * np.abs(gwo).max())


def test_wls_fp32_close_to_fp64(gpu_ready):
    torch = _torch()
    from cuphu_amd.autograd import WLSSolve

    w, phi, _ = _case(64, 80, seed=3)
    pt = torch.tensor(phi, device="cuda")
    u32 = WLSSolve.apply(torch.tensor(w, dtype=torch.float32, device="cuda"), pt, 4000, 1e-6)
    uo, _ = O.wls_solve_f64(w, phi.astype(np.float64))
    assert u32.dtype == torch.float32
    assert np.abs(u32.cpu().numpy() - uo).max() < 2e-3 * max(1.0, np.abs(uo).max())


def test_gradcheck_128_fp64(gpu_ready):
    """torch.autograd.gradcheck (fast mode: random projections of the
    Jacobian against central finite differences) on a 128x128 instance."""
    torch = _torch()
    from cuphu_amd.autograd import WLSSolve

    w, phi, _ = _case(128, 128, seed=11)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    pt = torch.tensor(phi, device="cuda")
    fn = lambda ww: WLSSolve.apply(ww, pt, 20000, 1e-13)  # noqa: E731
    # eps 1e-4 keeps the solve's own error (~1e-9 in u) well below the
    # difference quotient's resolution; a wrong adjoint is off by O(1)
    assert torch.autograd.gradcheck(fn, (wt,), eps=1e-4, atol=1e-4, rtol=1e-3, fast_mode=True)


def test_mapping_parameters_gradient_matches_finite_differences(gpu_ready):
    """loss = mean((u - truth)^2) with w = a c^b / (1 - c^2 + eps): gradients
    in a, b, eps by autograd + the adjoint solve vs central differences."""
    torch = _torch()
    from cuphu_amd.autograd import coherence_weights, wls_solve

    s = synth.make_scene(96, 112, nbumps=2, npatches=1, nlooks=10.0, seed=5, patch=20)
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    pt = torch.tensor(ph, device="cuda")
    ct = torch.tensor(s.corr, dtype=torch.float64, device="cuda")
    truth = torch.tensor(s.truth - s.truth.mean(), dtype=torch.float64, device="cuda")

    def loss_of(params):
        a, b, e = params
        w = coherence_weights(ct, a, b, e)
        u = wls_solve(w, pt, 20000, 1e-13)
        return torch.mean((u - truth) ** 2)

    p0 = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for v in (1.3, 2.0, 0.05)]
    loss = loss_of(p0)
    loss.backward()
    scale = max(abs(float(q.grad)) for q in p0)
    for k in range(3):
        h = 1e-3 * max(0.05, abs(float(p0[k])))
        pp = [q.detach().clone() for q in p0]
        pm = [q.detach().clone() for q in p0]
        pp[k] += h
        pm[k] -= h
        fd = (float(loss_of(pp)) - float(loss_of(pm))) / (2 * h)
        assert float(p0[k].grad) == pytest.approx(fd, rel=1e-3, abs=1e-4 * scale), k

"""Differentiable weighted-least-squares unwrapping (torch.autograd).

``WLSSolve`` solves the reference's weighted normal equations with fixed
arc weights w,

    L(w) u = b(w, phi),   L(w) = sum_a w_a (e_h - e_t)(e_h - e_t)^T + eps I,
    b(w, phi) = sum_a w_a g_a (e_h - e_t),  g_a = wrap(phi_h - phi_t),

(laplace_assemble / laplace_matvec, src/cxx/cuphu/cuphu_laplace.cu:88-190)
with the fixed-weight PCG of the Laplace path (no IRLS, no anchor, no
congruence) and returns u with its mean removed.  Its backward pass is the
adjoint: one more PCG solve with the same SPD operator,

    L(w) lambda = P dLoss/du   (P removes the mean),
    dLoss/dw_a = -(lambda_h - lambda_t) * ((u_h - u_t) - g_a),

(cuphu_wls_solve_gpu / cuphu_wls_adjoint_gpu).  Arc layout: the reference's
flat row-then-column order, narcs = (nrow-1)*ncol + nrow*(ncol-1).

``coherence_weights`` is the learnable coherence-to-weight mapping
w = a * c^b / (1 - c^2 + eps) evaluated on the arc coherence
c_arc = (c_t + c_h) / 2; its parameters get gradients from torch itself.
"""

from __future__ import annotations

import ctypes as C

import torch

from cuphu_amd import _capi

__all__ = ["WLSSolve", "wls_solve", "arc_coherence", "coherence_weights", "narcs"]


def narcs(nrow: int, ncol: int) -> int:
    return (nrow - 1) * ncol + nrow * (ncol - 1)


def _precision(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return 32
    if t.dtype == torch.float64:
        return 64
    raise TypeError(f"weights must be float32 or float64, got {t.dtype}")


def _check(w: torch.Tensor, phase: torch.Tensor):
    if not (w.is_cuda and phase.is_cuda):
        raise ValueError("w and phase must be GPU tensors")
    if phase.dtype != torch.float32 or phase.dim() != 2:
        raise TypeError("phase must be a 2-D float32 wrapped phase")
    nrow, ncol = phase.shape
    if nrow < 2 or ncol < 2:
        raise ValueError("phase must be at least 2x2")
    if w.dim() != 1 or w.numel() != narcs(nrow, ncol):
        raise ValueError(f"w must hold narcs = {narcs(nrow, ncol)} arc weights, got {tuple(w.shape)}")
    return nrow, ncol


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class WLSSolve(torch.autograd.Function):
    """u = WLSSolve.apply(w, phase, max_iter, tol).

    w: [narcs] float32 or float64 GPU tensor of arc weights (>= 0);
    phase: [nrow, ncol] float32 wrapped phase on the GPU.  u has w's dtype.
    Gradients flow to w only (phase is data)."""

    @staticmethod
    def forward(ctx, w, phase, max_iter=4000, tol=1e-8):
        nrow, ncol = _check(w, phase)
        w = w.contiguous()
        phase = phase.contiguous()
        u = torch.empty((nrow, ncol), dtype=w.dtype, device=w.device)
        it, rel = C.c_int(), C.c_double()
        _capi.check(_capi.lib.cuphu_wls_solve_gpu(_capi.ptr(w), _capi.ptr(phase), nrow, ncol, _precision(w),
                                                  int(max_iter), float(tol), _capi.ptr(u), C.byref(it),
                                                  C.byref(rel), _stream()), "cuphu_wls_solve_gpu")
        ctx.save_for_backward(w, phase, u)
        ctx.max_iter, ctx.tol = int(max_iter), float(tol)
        ctx.forward_iters, ctx.forward_relres = it.value, rel.value
        return u

    @staticmethod
    def backward(ctx, gu):
        w, phase, u = ctx.saved_tensors
        nrow, ncol = phase.shape
        gu = gu.contiguous().to(w.dtype)
        gw = torch.empty_like(w)
        it, rel = C.c_int(), C.c_double()
        _capi.check(_capi.lib.cuphu_wls_adjoint_gpu(_capi.ptr(w), _capi.ptr(phase), _capi.ptr(u), _capi.ptr(gu),
                                                    nrow, ncol, _precision(w), ctx.max_iter, ctx.tol,
                                                    _capi.ptr(gw), None, C.byref(it), C.byref(rel), _stream()),
                    "cuphu_wls_adjoint_gpu")
        return gw, None, None, None


def wls_solve(w: torch.Tensor, phase: torch.Tensor, max_iter: int = 4000, tol: float = 1e-8) -> torch.Tensor:
    """Functional form of WLSSolve.apply."""
    return WLSSolve.apply(w, phase, max_iter, tol)


def arc_coherence(corr: torch.Tensor) -> torch.Tensor:
    """Mean coherence of the two pixels of every arc, flat row-then-column
    arc order ([narcs])."""
    rows = 0.5 * (corr[:-1, :] + corr[1:, :])
    cols = 0.5 * (corr[:, :-1] + corr[:, 1:])
    return torch.cat([rows.reshape(-1), cols.reshape(-1)])


def coherence_weights(corr: torch.Tensor, a, b, eps) -> torch.Tensor:
    """w = a * c^b / (1 - c^2 + eps) on the arc coherence (differentiable in
    a, b, eps and corr)."""
    c = arc_coherence(corr)
    return a * torch.pow(c, b) / (1.0 - c * c + eps)

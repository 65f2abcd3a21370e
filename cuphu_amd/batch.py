"""Many interferograms in one launch sequence.

``unwrap_batch`` unwraps a list of independent interferograms (any sizes,
each with its own nlooks) with ``cuphu_laplace_batch_gpu``: one ingest, one
cost pass, one PCG/IRLS solve over all instances with per-instance
convergence masks, anchor + congruence and connected components.  Each
instance is unwrapped exactly as ``unwrap(..., init='laplace', ntiles=(1, 1),
reference_pixel=None)`` would unwrap it alone (the batch and solo results
are bit-identical; tests/test_gpu_batch.py).

The reference has no batched entry point; the per-instance semantics are
those of its single-tile Laplace path (src/cxx/cuphu/cuphu_solver.cpp:370-510,
src/cxx/cuphu/cuphu_laplace.cu:351-552).
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from cuphu_amd import _capi

__all__ = ["unwrap_batch"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("unwrap_batch needs a visible GPU")
    return torch


def _as_device(torch, a, dtype, name, i):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            a = a.cuda()
        if a.dtype != dtype:
            raise TypeError(f"{name}[{i}] must be {dtype}, got {a.dtype}")
        return a.contiguous()
    arr = np.ascontiguousarray(a)
    return torch.from_numpy(arr).to(device="cuda", dtype=dtype)


def unwrap_batch(
    igrams: Sequence,
    corrs: Sequence,
    nlooks,
    cost: str = "smooth",
    *,
    masks: Sequence | None = None,
    u0: Sequence | None = None,
    conncomp: bool = True,
    min_conncomp_frac: float | None = None,
    laplace_options: dict | None = None,
    return_stats: bool = False,
):
    """Unwrap ``len(igrams)`` interferograms together.

    igrams: complex64 arrays [nrow_i, ncol_i] (numpy, or torch tensors on the
    GPU); a float32 array is taken as an already-wrapped phase in [0, 2pi).
    corrs: float32 coherence of the same shapes.  nlooks: one float or one per
    instance.  masks / u0 (optional): uint8 validity masks / float32 warm
    starts per instance (entries may be None).  laplace_options: PCG/IRLS
    options as for ``unwrap`` (n_irls, delta_max, delta_min, max_iter, tol,
    solver, precision, rr_interval, use_graph, group_bytes_mb).

    Returns (unw list, conncomp list) -- numpy arrays when the inputs were
    numpy, GPU tensors otherwise -- plus the per-instance solve statistics
    when return_stats is set.  conncomp entries are None when conncomp is
    False.
    """
    torch = _torch()
    n = len(igrams)
    if n == 0:
        return ([], [], []) if return_stats else ([], [])
    if len(corrs) != n:
        raise ValueError("igrams and corrs must have the same length")
    if np.ndim(nlooks) == 0:
        nlooks = [float(nlooks)] * n
    if len(nlooks) != n:
        raise ValueError("nlooks must be a scalar or one value per instance")
    if cost not in _capi.COST:
        raise ValueError(f"cost must be one of {sorted(_capi.COST)}")
    numpy_out = not isinstance(igrams[0], torch.Tensor)

    insts = (_capi.CuPhuInstance * n)()
    keep, outs, ccs = [], [], []
    for i in range(n):
        ig = igrams[i]
        is_phase = (ig.dtype == np.float32) if not isinstance(ig, torch.Tensor) else ig.dtype == torch.float32
        if is_phase:
            dph = _as_device(torch, ig, torch.float32, "igrams", i)
            dig = None
            nr, nc = dph.shape
        else:
            dig = _as_device(torch, ig, torch.complex64, "igrams", i)
            nr, nc = dig.shape
            dph = None
        if nr < 1 or nc < 1:
            raise ValueError(f"instance {i} is empty")
        dc = _as_device(torch, corrs[i], torch.float32, "corrs", i)
        if tuple(dc.shape) != (nr, nc):
            raise ValueError(f"corrs[{i}] shape {tuple(dc.shape)} != {(nr, nc)}")
        dm = _as_device(torch, masks[i], torch.uint8, "masks", i) if masks is not None else None
        du0 = _as_device(torch, u0[i], torch.float32, "u0", i) if u0 is not None else None
        for t, nm in ((dm, "masks"), (du0, "u0")):
            if t is not None and tuple(t.shape) != (nr, nc):
                raise ValueError(f"{nm}[{i}] shape {tuple(t.shape)} != {(nr, nc)}")
        du = torch.empty((nr, nc), dtype=torch.float32, device="cuda")
        dl = torch.empty((nr, nc), dtype=torch.int32, device="cuda") if conncomp else None
        keep += [dig, dph, dc, dm, du0]
        insts[i] = _capi.CuPhuInstance(
            nr, nc, _capi.ptr(dig.view(torch.float32) if dig is not None else None), _capi.ptr(dph),
            _capi.ptr(dc), _capi.ptr(dm), _capi.ptr(du0), float(nlooks[i]), _capi.ptr(du), _capi.ptr(dl))
        outs.append(du)
        ccs.append(dl)

    prm = _capi.default_params()
    if min_conncomp_frac is not None:
        prm.minconncompfrac = float(min_conncomp_frac)
    opts = _capi.default_options(**(laplace_options or {}))
    stats = (_capi.CuPhuSolveStats * n)()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check(_capi.lib.cuphu_laplace_batch_gpu(insts, n, _capi.COST[cost], C.byref(prm), C.byref(opts), stats,
                                                  stream), "cuphu_laplace_batch_gpu")
    ccs = [c.view(torch.int32) if c is not None else None for c in ccs]
    if numpy_out:
        torch.cuda.current_stream().synchronize()
        outs = [u.cpu().numpy() for u in outs]
        ccs = [c.cpu().numpy().view(np.uint32) if c is not None else None for c in ccs]
    del keep
    if return_stats:
        return outs, ccs, [stats[i].as_dict() for i in range(n)]
    return outs, ccs

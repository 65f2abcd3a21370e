"""Helpers for GPU tests: device arrays through torch (allocation and
streams only) and thin wrappers over the C-ABI entry points."""

import ctypes as C

import numpy as np

from cuphu_amd import _capi


def torch_mod():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("torch sees no GPU")
    return torch


def dev(a, dtype=None):
    """numpy -> contiguous device tensor."""
    torch = torch_mod()
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype))
    return torch.from_numpy(a).to("cuda")


def empty(shape, dtype):
    torch = torch_mod()
    return torch.empty(shape, dtype=dtype, device="cuda")


def host(t):
    torch = torch_mod()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def stream():
    torch = torch_mod()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def narcs(nrow, ncol):
    return (nrow - 1) * ncol + nrow * (ncol - 1)


def batch_solve(scenes, opts=None, masks=None, want_cc=True, cost=0, params=None, phases=None):
    """Run cuphu_laplace_batch_gpu on a list of synth scenes; returns
    (list of unw, list of cc, stats list)."""
    torch = torch_mod()
    n = len(scenes)
    insts = (_capi.CuPhuInstance * n)()
    keep = []
    outs, ccs = [], []
    for i, s in enumerate(scenes):
        nr, nc = s.corr.shape
        if phases is not None:
            dph = dev(phases[i], np.float32)
            dig = None
        else:
            dig = dev(s.igram.view(np.float32).reshape(nr, nc * 2))
            dph = None
        dc = dev(s.corr, np.float32)
        dm = dev(masks[i], np.uint8) if masks is not None and masks[i] is not None else None
        du = empty((nr, nc), torch.float32)
        dl = empty((nr, nc), torch.int32) if want_cc else None
        keep += [dig, dph, dc, dm, du, dl]
        insts[i] = _capi.CuPhuInstance(nr, nc, p(dig), p(dph), p(dc), p(dm), None, float(s.nlooks), p(du), p(dl))
        outs.append(du)
        ccs.append(dl)
    prm = params if params is not None else _capi.default_params()
    o = opts if opts is not None else _capi.default_options()
    stats = (_capi.CuPhuSolveStats * n)()
    rc = _capi.lib.cuphu_laplace_batch_gpu(insts, n, cost, C.byref(prm), C.byref(o), stats, stream())
    _capi.check(rc, "cuphu_laplace_batch_gpu")
    torch.cuda.synchronize()
    u = [host(t) for t in outs]
    c = [host(t).view(np.uint32) if t is not None else None for t in ccs]
    return u, c, [stats[i].as_dict() for i in range(n)]


def sigsq_from_mask(ph, corr, nlooks, mask=None, kperp=7, kpar=7, defo=False):
    """Oracle sigsq with the reference's mask-derived arc weights."""
    import oracle as O

    w = None
    if mask is not None:
        m = mask.astype(bool)
        w = np.concatenate([(m[:-1, :] & m[1:, :]).ravel(), (m[:, :-1] & m[:, 1:]).ravel()]).astype(np.int16)
    return O.costs(ph, corr, nlooks, weights=w, kperp=kperp, kpar=kpar, defo=defo)

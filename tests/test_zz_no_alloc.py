"""cuphu_laplace_unwrap_gpu_noalloc (no counterpart in the reference; for
callers that cannot allocate) against cuphu_laplace_unwrap_gpu, through the
C ABI: the same unwrapped phase bit for bit on consecutive calls of
decreasing size (the fixed scratch keeps the previous call's contents),
on solves that converge inside a chunk or stop at a cap that is not a
multiple of it (the allocating entry pipelines its chunks, this one does
not), up to the largest instance it holds; no change in free device memory
across its calls; and a clear error, before anything runs, for an instance
beyond CUPHU_NOALLOC_MAX_ELEMENTS."""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from cuphu_amd import _capi, synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
MAX_EL = 1048576     # CUPHU_NOALLOC_MAX_ELEMENTS


def _hip():
    rt = C.CDLL(None)
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipFree.restype = C.c_int
    rt.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    rt.hipMemGetInfo.restype = C.c_int
    return rt


def _free_bytes():
    G.torch_mod().cuda.synchronize()
    free, total = C.c_size_t(), C.c_size_t()
    assert _hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


class _Case:
    """Device phase, device smooth costs and two output buffers."""

    def __init__(self, nrow, ncol, seed):
        s = synth.make_scene(nrow, ncol, nbumps=3, npatches=1, nlooks=10.0, seed=seed, patch=min(nrow, ncol) // 4)
        self.shape = (nrow, ncol)
        self.ph = O.wrapped_phase(s.igram, correctly_rounded=True)
        self.dph = G.dev(self.ph, np.float32)
        dco = G.dev(s.corr, np.float32)
        prm = _capi.default_params()
        prm.nlooks = prm.ncorrlooks = 10.0
        self.dcost = C.c_void_p()
        _capi.check(_capi.lib.cuphu_build_smooth_costs_gpu(G.p(self.dph), G.p(dco), None, None, nrow, ncol,
                                                           C.byref(prm), 7, 7, C.byref(self.dcost), G.stream()),
                    "smooth costs")
        torch = G.torch_mod()
        self.u_ref = G.empty(self.shape, torch.float32)
        self.u_na = G.empty(self.shape, torch.float32)
        self.u_na.fill_(np.nan)

    def run(self, noalloc, max_iter=300, tol=1e-6):
        fn = _capi.lib.cuphu_laplace_unwrap_gpu_noalloc if noalloc else _capi.lib.cuphu_laplace_unwrap_gpu
        out = self.u_na if noalloc else self.u_ref
        return fn(self.dcost, G.p(self.dph), self.shape[0], self.shape[1], 200.0, max_iter, C.c_float(tol), 0,
                  G.p(out), G.stream())

    def close(self):
        G.torch_mod().cuda.synchronize()
        assert _hip().hipFree(self.dcost.value) == 0


@pytest.mark.parametrize("merge", ["1", "2"])
def test_noalloc_equals_allocating_entry_on_two_calls(gpu_ready, monkeypatch, merge):
    """merge 2: the 517-column instance (512 + 64 work columns) runs its last
    64 columns in 64 x 64 tiles, whose tile list also lives in the scratch."""
    monkeypatch.setenv("CUPHU_N_IRLS", "3")
    monkeypatch.setenv("CUPHU_LAPLACE_MERGE", merge)
    big, small = _Case(300, 517, 61), _Case(131, 170, 62)
    try:
        for c in (big, small):
            _capi.check(c.run(False), "cuphu_laplace_unwrap_gpu")
        free0 = _free_bytes()
        for c in (big, small):                       # second call smaller than the first
            _capi.check(c.run(True), "cuphu_laplace_unwrap_gpu_noalloc")
            u_ref, u_na = G.host(c.u_ref), G.host(c.u_na)
            np.testing.assert_array_equal(u_na, u_ref)
            k = (u_na.astype(np.float64) - c.ph) / (2 * np.pi)
            assert np.abs(k - np.rint(k)).max() < 1e-4 and u_na[0, 0] == pytest.approx(float(c.ph[0, 0]), abs=1e-4)
        free1 = _free_bytes()
        print(f"free device memory across the no-allocation calls: {free0} -> {free1}")
        assert free1 == free0
    finally:
        big.close()
        small.close()


@pytest.mark.parametrize("tol,max_iter", [(3e-2, 3000), (1e-2, 3000), (1e-6, 130)])
def test_noalloc_equals_allocating_entry_on_converging_and_odd_capped_solves(gpu_ready, monkeypatch, tol, max_iter):
    """The allocating entry pipelines its 50-iteration chunks (the next one is
    issued before the state after the current one is read); the
    no-allocation entry reads the state after every chunk first.  Solves that
    converge inside a chunk (the chunk issued after it is a no-op) and a cap
    that is not a multiple of the chunk must give the same phase bit for
    bit."""
    monkeypatch.setenv("CUPHU_N_IRLS", "2")
    c = _Case(300, 517, 65)
    try:
        _capi.check(c.run(False, max_iter=max_iter, tol=tol), "cuphu_laplace_unwrap_gpu")
        _capi.check(c.run(True, max_iter=max_iter, tol=tol), "cuphu_laplace_unwrap_gpu_noalloc")
        np.testing.assert_array_equal(G.host(c.u_na), G.host(c.u_ref))
    finally:
        c.close()


def test_noalloc_at_its_largest_instance(gpu_ready, monkeypatch):
    """1024 x 1024 is exactly CUPHU_NOALLOC_MAX_ELEMENTS row-pitched
    elements; a smaller call follows it."""
    monkeypatch.setenv("CUPHU_N_IRLS", "1")
    cases = [_Case(1024, 1024, 63), _Case(257, 300, 64)]
    try:
        for c in cases:
            _capi.check(c.run(False, max_iter=120), "cuphu_laplace_unwrap_gpu")
            _capi.check(c.run(True, max_iter=120), "cuphu_laplace_unwrap_gpu_noalloc")
            np.testing.assert_array_equal(G.host(c.u_na), G.host(c.u_ref))
    finally:
        for c in cases:
            c.close()


@pytest.mark.parametrize("shape", [(1025, 1024), (1, MAX_EL + 1), (4096, 300)])
def test_noalloc_rejects_instances_beyond_its_scratch(gpu_ready, shape):
    """Checked on the host before any work: the pointers are never read."""
    torch = G.torch_mod()
    dummy = G.empty((64,), torch.float32)
    rc = _capi.lib.cuphu_laplace_unwrap_gpu_noalloc(G.p(dummy), G.p(dummy), shape[0], shape[1], 200.0, 10,
                                                    C.c_float(1e-6), 0, G.p(dummy), G.stream())
    assert rc == -2
    msg = _capi.lib.cuphu_last_error().decode()
    assert "no-allocation" in msg and str(MAX_EL) in msg, msg

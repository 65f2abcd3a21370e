"""The reference-named device-pointer entry points, called through the C
ABI and checked against the oracle:

* cuphu_build_smooth_costs_gpu / cuphu_build_defo_costs_gpu (ref
  cuphu_cost.cu:338-351, :417-430): device phase, coherence and optional
  row/column arc weights in; the callee allocates the flat row-then-column
  cost array, which the caller frees with hipFree.
* cuphu_laplace_unwrap_gpu (ref cuphu_laplace.cu:351-363): smooth costs
  and phase in, the anchored congruent unwrapped phase out; the IRLS
  schedule comes from CUPHU_N_IRLS / CUPHU_IRLS_DELTA_MAX /
  CUPHU_IRLS_DELTA_MIN, as in the reference.

Costs may differ from the oracle by +-1 at truncation boundaries (device
versus host FP64 pow); those cases are counted and bounded.  The rhopow
case with chosen constants has no boundary nearby and is exact.
"""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from cuphu_amd import _capi, synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def _hip():
    """hipMemcpy / hipFree of the process's one HIP runtime (loaded
    RTLD_GLOBAL before the library)."""
    rt = C.CDLL(None)
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipFree.restype = C.c_int
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipMemcpy.restype = C.c_int
    return rt


def _device_costs(ph, corr, params, defo, weights=None, kperp=7, kpar=7):
    """Runs the device cost builder; returns the host copy [narcs, 2|4]."""
    nrow, ncol = ph.shape
    na = G.narcs(nrow, ncol)
    dph, dco = G.dev(ph, np.float32), G.dev(corr, np.float32)
    nra = (nrow - 1) * ncol
    dr = G.dev(weights[:nra], np.int16) if weights is not None else None
    dc = G.dev(weights[nra:], np.int16) if weights is not None else None
    out = C.c_void_p()
    fn = _capi.lib.cuphu_build_defo_costs_gpu if defo else _capi.lib.cuphu_build_smooth_costs_gpu
    _capi.check(fn(G.p(dph), G.p(dco), G.p(dr), G.p(dc), nrow, ncol, C.byref(params), kperp, kpar, C.byref(out),
                   G.stream()), "device costs")
    G.torch_mod().cuda.synchronize()
    ne = 4 if defo else 2
    host = np.empty((na, ne), np.int16)
    rt = _hip()
    assert rt.hipMemcpy(host.ctypes.data, out.value, host.nbytes, 2) == 0     # device to host
    assert rt.hipFree(out.value) == 0
    return host


def _params(nlooks, **over):
    prm = _capi.default_params()
    prm.nlooks = prm.ncorrlooks = nlooks
    for k, v in over.items():
        setattr(prm, k, v)
    return prm


@pytest.mark.parametrize("defo", [False, True])
@pytest.mark.parametrize("nlooks,kperp,kpar,masked", [(5.0, 7, 7, False), (23.8, 3, 5, True), (10.0, 7, 7, True)])
def test_device_cost_builders_match_oracle(gpu_ready, defo, nlooks, kperp, kpar, masked):
    s = synth.make_scene(83, 121, nbumps=3, npatches=1, nlooks=nlooks, seed=31, patch=25)
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    nrow, ncol = ph.shape
    w = None
    if masked:
        w = np.ones(G.narcs(nrow, ncol), np.int16)
        w[np.random.default_rng(5).random(w.size) < 0.1] = 0
    got = _device_costs(ph, s.corr, _params(nlooks), defo, w, kperp, kpar)
    ref = O.costs(ph, s.corr, nlooks, weights=w, kperp=kperp, kpar=kpar, defo=defo)
    d = np.abs(got.astype(np.int32) - ref)
    print(f"device costs defo={defo} nlooks={nlooks}: {(d > 0).sum()} +-1 cases of {d.size}")
    assert d.max() <= 1 and (d > 0).sum() <= 3
    if w is not None:
        assert (got[w == 0, 1] == 32000).all()


@pytest.mark.parametrize("defo", [False, True])
@pytest.mark.parametrize("ncorrlooks,cstd2,cstd3,rho,expect", [
    (float(np.exp(2.0)), 1.0, 0.0, 0.5, 36),      # rhopow = ln(e^2) = 2
    (10.0, 0.0, 0.1, 0.5, 53),                    # rhopow = 0.1 * 10 = 1
    (10.0, 0.35, 0.06, 0.6, 28),                  # defaults with cstd1 = 0.4: rhopow 2.2059
])
def test_device_costs_hand_evaluated_rhopow(gpu_ready, defo, ncorrlooks, cstd2, cstd3, rho, expect):
    """sigsq = ((1 - rho)^rhopow / 6 + 0.05) * 200^2 / 100, evaluated by
    hand, on a flat phase with uniform coherence (every arc the same)."""
    cstd1 = 0.4 if expect == 28 else 0.0
    ph = np.zeros((9, 14), np.float32)
    corr = np.full((9, 14), rho, np.float32)
    got = _device_costs(ph, corr, _params(ncorrlooks, cstd1=cstd1, cstd2=cstd2, cstd3=cstd3), defo)
    assert (got[:, 1] == expect).all()
    assert (got[:, 0] == 0).all()                 # zero gradients: zero offsets


@pytest.mark.parametrize("n_irls", [0, 3])
def test_laplace_unwrap_gpu_matches_oracle(gpu_ready, monkeypatch, n_irls):
    monkeypatch.setenv("CUPHU_N_IRLS", str(n_irls))
    monkeypatch.setenv("CUPHU_IRLS_DELTA_MAX", "0.5")
    monkeypatch.setenv("CUPHU_IRLS_DELTA_MIN", "0.05")
    s = synth.make_scene(140, 170, nbumps=3, npatches=1, nlooks=10.0, seed=41, patch=30)
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    nrow, ncol = ph.shape
    prm = _params(10.0)
    dph, dco = G.dev(ph, np.float32), G.dev(s.corr, np.float32)
    dcost = C.c_void_p()
    _capi.check(_capi.lib.cuphu_build_smooth_costs_gpu(G.p(dph), G.p(dco), None, None, nrow, ncol, C.byref(prm), 7, 7,
                                                       C.byref(dcost), G.stream()), "smooth costs")
    du = G.empty((nrow, ncol), G.torch_mod().float32)
    _capi.check(_capi.lib.cuphu_laplace_unwrap_gpu(dcost, G.p(dph), nrow, ncol, 200.0, 800, C.c_float(1e-6), 0,
                                                   G.p(du), G.stream()), "cuphu_laplace_unwrap_gpu")
    u = G.host(du)
    assert _hip().hipFree(dcost.value) == 0
    sig = O.costs(ph, s.corr, 10.0)[:, 1]
    opts = dict(n_irls=n_irls, delta_max=0.5, delta_min=0.05, max_iter=800, tol=1e-6)
    ref, _ = O.pcg_irls_f64(sig, ph, **opts)
    ka = np.rint((u.astype(np.float64) - ph) / TWOPI)
    kb = np.rint((ref - ph) / TWOPI)
    agree = float((ka == kb).mean())
    print(f"cuphu_laplace_unwrap_gpu n_irls={n_irls}: agreement with the FP64 oracle {agree:.6f}")
    assert agree >= 0.999
    assert u[0, 0] == pytest.approx(float(ph[0, 0]), abs=1e-4)      # anchor (cuphu_laplace.cu:534-539)
    k = (u.astype(np.float64) - ph) / TWOPI                          # congruence (cuphu_laplace.cu:256-266)
    assert np.abs(k - np.rint(k)).max() < 1e-4

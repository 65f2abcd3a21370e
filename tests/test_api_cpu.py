"""Host-side behaviour and the C-ABI boundary (CPU only, no GPU calls).

Covers the reference's pure-Python tests (test_unwrap.py:44-80,
test_auto_tiling.py, test_mask_pad_distance.py:11-25, test_looks.py), the
struct layout of include/cuphu.h against the ctypes mirror, the exported
symbols, and the host-only C-ABI functions (tile geometry, stitching)
against the oracle.
"""

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi
from cuphu_amd._check import check_cost_mode, check_init_method
from cuphu_amd._unwrap import _auto_ntiles_by_target_size, _dilate_mask

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cuphu.h")


@pytest.fixture()
def small_igram():
    r = np.arange(50, dtype=np.float32)
    c = np.arange(60, dtype=np.float32)
    return np.exp(1j * (0.1 * r[:, None] + 0.2 * c[None, :])).astype(np.complex64)


@pytest.fixture()
def small_corr():
    return np.full((50, 60), 0.9, dtype=np.float32)


# ── validation (test_unwrap.py:44-80) ───────────────────────────────────────
def test_invalid_cost_mode():
    with pytest.raises(ValueError, match="cost must be"):
        check_cost_mode("bilinear")


def test_invalid_init_method():
    with pytest.raises(ValueError, match="init must be"):
        check_init_method("dijkstra")


def test_igram_not_2d(small_corr):
    with pytest.raises(ValueError, match="2-D"):
        cuphu.unwrap(np.ones((2, 50, 60), dtype=np.complex64), small_corr, 10.0)


def test_shape_mismatch(small_igram):
    with pytest.raises(ValueError, match="shape"):
        cuphu.unwrap(small_igram, np.ones((10, 10), dtype=np.float32), 10.0)


def test_nlooks_too_small(small_igram, small_corr):
    with pytest.raises(ValueError, match="nlooks"):
        cuphu.unwrap(small_igram, small_corr, 0.5)


def test_wrong_igram_dtype(small_corr):
    with pytest.raises(TypeError, match="complex"):
        cuphu.unwrap(np.ones((50, 60), dtype=np.float32), small_corr, 10.0)


def test_wrong_corr_dtype(small_igram):
    with pytest.raises(TypeError, match="floating"):
        cuphu.unwrap(small_igram, np.ones((50, 60), dtype=np.int16), 10.0)


def test_wrong_mask_dtype(small_igram, small_corr):
    with pytest.raises(TypeError, match="8-bit"):
        cuphu.unwrap(small_igram, small_corr, 10.0, mask=np.ones((50, 60), np.float32))


# ── auto tiling (test_auto_tiling.py) ───────────────────────────────────────
def test_auto_ntiles_targets_edge_length():
    ntr, ntc = _auto_ntiles_by_target_size(18240, 13162, 1024, 64, 64)
    assert 700 < 18240 / ntr < 2000 and 700 < 13162 / ntc < 2000
    assert ntr >= ntc


def test_auto_ntiles_small_and_square():
    assert _auto_ntiles_by_target_size(200, 200, 1024, 64, 64) == (1, 1)
    ntr, ntc = _auto_ntiles_by_target_size(4096, 4096, 1024, 64, 64)
    assert ntr == ntc
    assert _auto_ntiles_by_target_size(2048, 2048, 1024, 64, 64) == (2, 2)
    assert _auto_ntiles_by_target_size(1024, 1024, 1024, 64, 64) == (1, 1)


# ── mask dilation (test_mask_pad_distance.py:11-25) ─────────────────────────
def test_dilate_mask_grows_by_exact_distance():
    m = np.zeros((21, 21), dtype=bool)
    m[10, 10] = True
    grown = _dilate_mask(m, 3)
    ys, xs = np.where(grown)
    assert (np.abs(ys - 10) + np.abs(xs - 10)).max() == 3
    assert grown.sum() == 1 + 4 + 8 + 12


@pytest.mark.parametrize("pixels", [1, 2, 5])
def test_dilate_mask_matches_brute_force_manhattan_distance(pixels):
    """Every pixel within Manhattan distance `pixels` of a valid pixel, by
    brute force over all valid pixels, on random sparse masks (the edges
    included) and on an empty one."""
    rng = np.random.default_rng(pixels)
    for m in (rng.random((17, 23)) > 0.93, np.zeros((6, 7), bool)):
        ys, xs = np.nonzero(m)
        rr, cc = np.mgrid[0:m.shape[0], 0:m.shape[1]]
        if ys.size:
            d = np.min(np.abs(rr[..., None] - ys) + np.abs(cc[..., None] - xs), axis=-1)
            expect = d <= pixels
        else:
            expect = np.zeros_like(m)
        np.testing.assert_array_equal(_dilate_mask(m, pixels), expect)


def test_dilate_mask_zero_is_noop():
    m = np.random.default_rng(0).random((30, 30)) > 0.5
    np.testing.assert_array_equal(_dilate_mask(m, 0), m)


def test_effective_looks_known_answer():
    assert cuphu.get_effective_looks(13, 16, 2.33, 2.75, 7.5, 6.0) == 29.616888888888894


def test_effective_looks_reference_properties():
    """test_looks.py:17-45: below the naive product when oversampled, equal
    to it at Nyquist spacing, linear in the number of looks."""
    kw = dict(range_spacing=2.33, azimuth_spacing=2.75, range_resolution=7.5, azimuth_resolution=6.0)
    assert cuphu.get_effective_looks(range_looks=13, azimuth_looks=16, **kw) < 13 * 16
    assert cuphu.get_effective_looks(range_looks=13, azimuth_looks=16, range_spacing=3.0, azimuth_spacing=4.0,
                                     range_resolution=3.0, azimuth_resolution=4.0) == pytest.approx(13 * 16)
    n1 = cuphu.get_effective_looks(range_looks=1, azimuth_looks=1, **kw)
    assert cuphu.get_effective_looks(range_looks=2, azimuth_looks=3, **kw) == pytest.approx(6 * n1)


# ── the C ABI boundary ──────────────────────────────────────────────────────
def _header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cuphu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    names = _header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(_capi.lib, n), n
    assert set(names) == set(_capi.EXPORTED)


def test_struct_layout_matches_header():
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "cuphu.h"
#define P(T) printf(#T " %zu\n", sizeof(T));
#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  P(CuPhuParams) P(CuPhuTileParams) P(CuPhuBridgeParams) P(CuPhuResult) P(CuPhuLaplaceOptions)
  P(CuPhuSolveStats) P(CuPhuInstance) P(CuPhuSmoothCost) P(CuPhuCost)
  F(CuPhuParams, sigsqshortmin) F(CuPhuParams, transmitmode) F(CuPhuParams, shadow)
  F(CuPhuParams, kperpdpsi) F(CuPhuParams, minconncompfrac) F(CuPhuParams, maxncomps)
  F(CuPhuLaplaceOptions, tol) F(CuPhuLaplaceOptions, verbose) F(CuPhuLaplaceOptions, autotune) F(CuPhuSolveStats, relres)
  F(CuPhuSolveStats, pass_status) F(CuPhuInstance, nlooks) F(CuPhuInstance, d_conncomp)
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    ct = {
        "CuPhuParams": _capi.CuPhuParams, "CuPhuTileParams": _capi.CuPhuTileParams,
        "CuPhuBridgeParams": _capi.CuPhuBridgeParams, "CuPhuResult": _capi.CuPhuResult,
        "CuPhuLaplaceOptions": _capi.CuPhuLaplaceOptions, "CuPhuSolveStats": _capi.CuPhuSolveStats,
        "CuPhuInstance": _capi.CuPhuInstance,
    }
    for k, v in ct.items():
        assert int(got[k]) == C.sizeof(v), k
    assert int(got["CuPhuSmoothCost"]) == 4 and int(got["CuPhuCost"]) == 8
    for key, val in got.items():
        if "." in key:
            s, f = key.split(".")
            f = "lambda_" if f == "lambda" else f
            assert getattr(ct[s], f).offset == int(val), key


def test_default_parameter_tables():
    p = _capi.default_params()
    assert (p.rhosconst1, p.rhosconst2, p.cstd1, p.cstd2, p.cstd3) == (1.3, 0.14, 0.4, 0.35, 0.06)
    assert (p.defothreshfactor, p.sigsqcorr, p.costscale, p.nshortcycle, p.sigsqshortmin) == (1.2, 0.05, 100, 200, 1)
    assert (p.kperpdpsi, p.kpardpsi, p.defomax, p.minconncompfrac, p.maxncomps) == (7, 7, 1.2, 0.01, 32)
    t = _capi.CuPhuTileParams()
    _capi.lib.cuphu_default_tile_params(C.byref(t))
    assert (t.ntilerow, t.ntilecol, t.tilecostthresh, t.minregionsize, t.nproc) == (1, 1, 500, 100, 1)
    assert t.laplace_neighbor_feedback_feather == 200 and t.single_tile_reoptimize == 0
    b = _capi.CuPhuBridgeParams()
    _capi.lib.cuphu_default_bridge_params(C.byref(b))
    assert (b.enabled, b.radius, b.min_num_pixel, b.erosion_size, b.max_boundary_samples) == (0, 500, 14, 2, 4096)
    o = _capi.default_options()
    assert (o.n_irls, o.max_iter, o.rr_interval) == (16, 1000, 50)
    assert o.tol == pytest.approx(1e-3) and o.delta_max == 0.5 and o.delta_min == pytest.approx(0.01)


def test_unsupported_and_invalid_requests_return_status():
    n = 8
    re_ = np.ones(n * n, np.float32)
    im_ = np.zeros(n * n, np.float32)
    corr = np.ones(n * n, np.float32)
    p = _capi.default_params()
    t = _capi.CuPhuTileParams()
    _capi.lib.cuphu_default_tile_params(C.byref(t))
    res = _capi.CuPhuResult()
    rc = _capi.lib.cuphu_unwrap(re_.ctypes.data, im_.ctypes.data, corr.ctypes.data, None, None, n, n, 0,
                                _capi.INIT["mcf"], C.byref(p), C.byref(t), None, None, 0, C.byref(res))
    assert rc == _capi.ERR_UNSUPPORTED
    assert b"network-flow" in _capi.lib.cuphu_last_error()
    rc = _capi.lib.cuphu_unwrap(re_.ctypes.data, im_.ctypes.data, corr.ctypes.data, None, None, 0, n, 0,
                                _capi.INIT["laplace"], C.byref(p), C.byref(t), None, None, 0, C.byref(res))
    assert rc == _capi.ERR_INVALID
    rc = _capi.lib.cuphu_unwrap(None, None, corr.ctypes.data, None, None, n, n, 0, 2, C.byref(p), C.byref(t),
                                None, None, 0, C.byref(res))
    assert rc == _capi.ERR_INVALID
    t.ntilerow, t.single_tile_reoptimize = 2, 1   # a full-scene CPU TreeSolve: needs SNAPHU
    rc = _capi.lib.cuphu_unwrap(re_.ctypes.data, im_.ctypes.data, corr.ctypes.data, None, None, n, n, 0, 2,
                                C.byref(p), C.byref(t), None, None, 0, C.byref(res))
    assert rc == _capi.ERR_UNSUPPORTED
    assert b"single_tile_reoptimize" in _capi.lib.cuphu_last_error()


def test_mcf_init_raises_through_python(small_igram, small_corr):
    with pytest.raises(ValueError, match="network-flow"):
        cuphu.unwrap(small_igram, small_corr, 10.0, init="mcf")


# ── host-only C-ABI functions vs the oracle ─────────────────────────────────
def _lib_tile_axis(n, ts, ov, nt=1):
    f = (C.c_int * 4096)()
    s = (C.c_int * 4096)()
    k = _capi.check(_capi.lib.cuphu_tile_axis(n, ts, ov, nt, f, s, 4096), "tile_axis")
    return list(f[:k]), list(s[:k])


def test_library_tile_axis_matches_oracle():
    rng = np.random.default_rng(1)
    for _ in range(300):
        n = int(rng.integers(20, 20000))
        ov = int(rng.integers(0, 300))
        nt = int(rng.integers(1, 12))
        assert _lib_tile_axis(n, 0, ov, nt) == O.tile_axis(n, 0, ov, nt)
        ts = int(rng.integers(ov + 16, ov + 3000))
        assert _lib_tile_axis(n, ts, ov) == O.tile_axis(n, ts, ov)
    assert _lib_tile_axis(8192, 2048, 256) == ([0, 1792, 3584, 5376, 7168], [2048] * 4 + [1024])


@pytest.mark.parametrize("seed,ntr,ntc,conf", [(0, 2, 3, False), (1, 3, 3, True), (2, 1, 4, False)])
def test_library_stitch_matches_oracle(seed, ntr, ntc, conf):
    rng = np.random.default_rng(seed)
    nrow, ncol, ov = 41, 57, 6
    rf, rs = O.tile_axis(nrow, 0, ov, ntr)
    cf, cs = O.tile_axis(ncol, 0, ov, ntc)
    base = np.cumsum(rng.normal(0, 0.4, (nrow, ncol)), axis=1).astype(np.float32)
    corr = rng.uniform(0.0, 1.0, (nrow, ncol)).astype(np.float32)
    tu, tc, tr_, tf = [], [], [], []
    for a in range(ntr):
        for b in range(ntc):
            sl = (slice(rf[a], rf[a] + rs[a]), slice(cf[b], cf[b] + cs[b]))
            off = np.float32(rng.uniform(-20, 20))
            tu.append(np.ascontiguousarray(base[sl] + off + rng.normal(0, 0.05, (rs[a], cs[b])).astype(np.float32)))
            tc.append(rng.integers(0, 4, (rs[a], cs[b])).astype(np.uint32))
            tr_.append(np.ascontiguousarray(corr[sl]))
            tf.append((rng.uniform(size=(rs[a], cs[b])) > 0.1).astype(np.uint8) if conf else None)
    eu, ec, ek = O.stitch(nrow, ncol, rf, rs, cf, cs, ov, ov, tu, tc, tr_, tf if conf else None, laplace=True)
    nt = ntr * ntc
    P = lambda arrs: (C.c_void_p * nt)(*[a.ctypes.data if a is not None else None for a in arrs])  # noqa: E731
    iv = lambda v: (C.c_int * len(v))(*v)  # noqa: E731
    unw = np.empty((nrow, ncol), np.float32)
    cc = np.empty((nrow, ncol), np.uint32)
    tk = (C.c_double * nt)()
    _capi.check(_capi.lib.cuphu_stitch_tiles_host(
        nrow, ncol, ntr, ntc, iv(rf), iv(rs), iv(cf), iv(cs), ov, ov, 1, P(tu), P(tc), P(tr_),
        P(tf) if conf else None, unw.ctypes.data, cc.ctypes.data, tk), "stitch")
    np.testing.assert_array_equal(np.array(tk[:]), ek)
    np.testing.assert_array_equal(unw, eu)
    np.testing.assert_array_equal(cc, ec)

"""The 8192x8192 auto-tiled scene (25 ragged 2048-px tiles, 256-px overlap)
through the device-resident scene path at full size, checked by properties
that do not depend on size: congruence with the wrapped phase, the anchor
at pixel (0,0), label range, tile-stitch consistency in the overlaps, and
bitwise repeatability; and the reference tile formula's 1843-column tiles
bit-identical on 64- and 256-element row strides and in the merged and
plain iteration tilings; and a 16384x16384 scene
whose pooled planes pass 2^31 bytes.  (Full-size parity against the FP64
solve is bench.py's accuracy leg.)"""

import ctypes as C

import numpy as np
import pytest

from cuphu_amd import _capi, synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def test_scene8192_properties(gpu_ready):
    torch = G.torch_mod()
    s = synth.config_scene("scene8192")
    nrow, ncol = s.corr.shape
    dig = torch.from_numpy(s.igram.view(np.float32).reshape(nrow, 2 * ncol)).to("cuda")
    dcorr = torch.from_numpy(s.corr).to("cuda")
    prm = _capi.default_params()
    prm.nlooks = prm.ncorrlooks = s.nlooks
    opts = _capi.default_options(n_irls=1, max_iter=100, tol=1e-6)
    outs = []
    for _ in range(2):
        du = torch.empty((nrow, ncol), dtype=torch.float32, device="cuda")
        dl = torch.empty((nrow, ncol), dtype=torch.int32, device="cuda")
        stats = (_capi.CuPhuSolveStats * 64)()
        nt = C.c_int()
        _capi.check(_capi.lib.cuphu_unwrap_scene_gpu(G.p(dig), G.p(dcorr), None, None, nrow, ncol, s.nlooks,
                                                     _capi.COST["defo"], C.byref(prm), 2048, 256, 1, 1, C.byref(opts),
                                                     G.p(du), G.p(dl), stats, C.byref(nt), G.stream()), "scene")
        torch.cuda.synchronize()
        assert nt.value == 25
        outs.append((du.cpu().numpy(), dl.cpu().numpy().view(np.uint32)))
    (u, lab), (u2, lab2) = outs
    np.testing.assert_array_equal(u, u2)
    np.testing.assert_array_equal(lab, lab2)
    assert np.isfinite(u).all()
    ph = np.mod(np.angle(s.igram.astype(np.complex128)), TWOPI)
    # congruence up to the per-tile fractional stitch offsets: within each
    # trimmed tile region the cycle count is an integer plus that tile's offset
    k = (u.astype(np.float64) - ph) / TWOPI
    frac = k - np.rint(k)
    starts = [0, 1792, 3584, 5376, 7168]
    for r in starts:
        for c in starts:
            blk = frac[r + 300:r + 1500, c + 300:c + 1500]
            if blk.size:
                assert np.ptp(blk) < 1e-3 or np.ptp(np.mod(blk + 0.5, 1.0)) < 1e-3
    assert lab.max() <= prm.maxncomps * 25
    assert (lab > 0).mean() > 0.9


def test_scene8192_reference_tiling_strides_bit_identical(gpu_ready, monkeypatch):
    """The reference tile formula at 8192^2 (5x5 tiles of 1843 columns, the
    Python API's default tiling): row strides rounded to 256 elements (the
    fused solver's default; the iteration kernel covers the 1856 work
    columns of a 2048-element stride) give the same unwrapped phase and
    labels, bit for bit, as strides rounded to 64; and so does the plain
    256-column tiling against the merged one (the last 64 of the 1856 work
    columns in 64 x 64 tiles, the default)."""
    torch = G.torch_mod()
    s = synth.config_scene("scene8192")
    nrow, ncol = s.corr.shape
    dig = torch.from_numpy(s.igram.view(np.float32).reshape(nrow, 2 * ncol)).to("cuda")
    dcorr = torch.from_numpy(s.corr).to("cuda")
    prm = _capi.default_params()
    prm.nlooks = prm.ncorrlooks = s.nlooks
    opts = _capi.default_options(n_irls=1, max_iter=60, tol=1e-6)
    outs = {}
    for stride, merge in (("64", "1"), ("256", "1"), ("256", "0")):
        monkeypatch.setenv("CUPHU_LAPLACE_STRIDE", stride)
        monkeypatch.setenv("CUPHU_LAPLACE_MERGE", merge)
        du = torch.empty((nrow, ncol), dtype=torch.float32, device="cuda")
        dl = torch.empty((nrow, ncol), dtype=torch.int32, device="cuda")
        stats = (_capi.CuPhuSolveStats * 64)()
        nt = C.c_int()
        _capi.check(_capi.lib.cuphu_unwrap_scene_gpu(G.p(dig), G.p(dcorr), None, None, nrow, ncol, s.nlooks,
                                                     _capi.COST["defo"], C.byref(prm), 0, 256, 5, 5, C.byref(opts),
                                                     G.p(du), G.p(dl), stats, C.byref(nt), G.stream()), "scene")
        torch.cuda.synchronize()
        assert nt.value == 25
        outs[stride, merge] = (du.cpu().numpy(), dl.cpu().numpy(), [stats[i].as_dict()["iters"] for i in range(25)])
    ref = outs["256", "1"]
    for key in (("64", "1"), ("256", "0")):
        np.testing.assert_array_equal(outs[key][0], ref[0], err_msg=str(key))
        np.testing.assert_array_equal(outs[key][1], ref[1], err_msg=str(key))
        assert outs[key][2] == ref[2]
    assert np.isfinite(ref[0]).all()


def test_scene16384_beyond_2gb_plane_offsets(gpu_ready):
    """A 16384 x 16384 scene (81 tiles of 2048 with 256 overlap; the pooled
    planes hold 3.4 x 10^8 elements, so byte offsets into them pass 2^31):
    a ramp with 0.2 rad of noise and coherence 0.8 everywhere, one PCG
    solve of 20 iterations.  Checked by properties: finite, congruent with
    the wrapped phase up to one fractional offset per tile (measured below
    1e-6 of a cycle here), the anchor pixel, every pixel labelled.  A
    32768^2 scene (361 tiles, about 98 GB of device memory) ran the same way
    outside the test suite."""
    torch = G.torch_mod()
    n = 16384
    rng = np.random.default_rng(3)
    r = np.arange(n, dtype=np.float32)[:, None]
    c = np.arange(n, dtype=np.float32)[None, :]
    ph = (0.003 * r + 0.002 * c).astype(np.float32) + rng.normal(0, 0.2, (n, n)).astype(np.float32)
    ig = np.empty((n, n), np.complex64)
    ig.real = np.cos(ph)
    ig.imag = np.sin(ph)
    dig = torch.from_numpy(ig.view(np.float32).reshape(n, 2 * n)).to("cuda")
    dcorr = torch.full((n, n), 0.8, dtype=torch.float32, device="cuda")
    du = torch.empty((n, n), dtype=torch.float32, device="cuda")
    dl = torch.empty((n, n), dtype=torch.int32, device="cuda")
    prm = _capi.default_params()
    prm.nlooks = prm.ncorrlooks = 10.0
    opts = _capi.default_options(n_irls=0, max_iter=20, tol=1e-6)
    stats = (_capi.CuPhuSolveStats * 128)()
    nt = C.c_int()
    _capi.check(_capi.lib.cuphu_unwrap_scene_gpu(G.p(dig), G.p(dcorr), None, None, n, n, 10.0, _capi.COST["defo"],
                                                 C.byref(prm), 2048, 256, 0, 0, C.byref(opts), G.p(du), G.p(dl), stats,
                                                 C.byref(nt), G.stream()), "scene")
    torch.cuda.synchronize()
    assert nt.value == 81
    u = du.cpu().numpy()
    lab = dl.cpu().numpy().view(np.uint32)
    del dig, dcorr, du, dl
    assert np.isfinite(u).all()
    wr = np.mod(np.angle(ig.astype(np.complex128)), TWOPI)
    k = (u.astype(np.float64) - wr) / TWOPI
    frac = np.abs(k - np.rint(k))
    assert frac.max() < 1e-3
    assert u[0, 0] == pytest.approx(float(wr[0, 0]), abs=1e-4)
    assert lab.min() >= 1

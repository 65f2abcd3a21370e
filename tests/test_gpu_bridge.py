"""Phase bridging on the GPU (reference cuphu_solver.cpp:1000-1366,
cuphu_bridge_gpu.cu).

The reference's own hook tests (test_bridge.py) run through the
reference-named binding; every result (unw and the region labels) is
compared bit for bit with the oracle, on the reference's fixtures and on
random multi-region scenes over all ramp types, erosion sizes and masks;
unwrap(bridge=True) is checked against the oracle applied to the unbridged
result."""

import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi, synth
from cuphu_amd._ext import _cuphu_ext
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * np.pi
RAMPS = [None, "linear", "quadratic", "linear_range", "linear_azimuth", "quadratic_range", "quadratic_azimuth"]


def _bridge(unw, mask=None, **kw):
    """GPU hook and oracle; asserts equality and returns the GPU result."""
    args = dict(bridge_radius=500, bridge_min_num_pixel=14, bridge_erosion_size=2,
                bridge_max_boundary_samples=4096, bridge_ramp_type=None, bridge_ramp_max_num_sample=1000000)
    args.update(kw)
    u, cc = _cuphu_ext._bridge_apply_test(unw, mask=mask, **args)
    eu, ecc, nreg = O.bridge(unw, mask, radius=args["bridge_radius"], min_num_pixel=args["bridge_min_num_pixel"],
                             erosion_size=args["bridge_erosion_size"],
                             max_boundary_samples=args["bridge_max_boundary_samples"],
                             ramp_type=args["bridge_ramp_type"], ramp_max_num_sample=args["bridge_ramp_max_num_sample"])
    np.testing.assert_array_equal(u, eu)
    np.testing.assert_array_equal(cc, ecc)
    return u, cc, nreg


# ── the reference's hook tests ────────────────────────────────────────────────

def test_bridge_reference_cases(gpu_ready):
    nrow, ncol = 60, 90
    rng = np.random.default_rng(0)
    r, c = np.mgrid[0:nrow, 0:ncol]
    ramp = 0.02 * r + 0.015 * c
    off = np.zeros((nrow, ncol))
    off[(r >= 20) & (r < 39)] = 3.0 * TWO_PI
    off[r >= 40] = -2.0 * TWO_PI
    unw = (ramp + off).astype(np.float32) + rng.normal(0.0, 0.05, (nrow, ncol)).astype(np.float32)
    m = np.ones((nrow, ncol), np.uint8)
    m[19, :] = 0
    m[39, :] = 0
    u, cc, n = _bridge(unw, m, bridge_min_num_pixel=1, bridge_erosion_size=0)
    v = m == 1
    assert n == 3 and (u[v] - ramp[v]).std() < 0.3 and not np.allclose(u[v], unw[v])
    assert len(np.unique(cc[v])) >= 1

    unw = np.full((30, 30), 1.2345, np.float32)
    u, cc, _ = _bridge(unw)
    np.testing.assert_array_equal(u, unw)

    r, c = np.mgrid[0:40, 0:60]
    ramp = 1.0 + 0.02 * r + 0.01 * c
    unw = (ramp + np.where(c >= 30, 4.0 * TWO_PI, 0.0)).astype(np.float32)
    m = np.ones((40, 60), np.uint8)
    m[:, 28:32] = 0
    unw[:, 28:32] = 99.0
    u, _, _ = _bridge(unw, m, bridge_min_num_pixel=1, bridge_erosion_size=0)
    v = m == 1
    assert (u[v] - ramp[v]).std() < 0.3
    np.testing.assert_array_equal(u[:, 28:32], unw[:, 28:32])

    r, _ = np.mgrid[0:200, 0:10]
    ramp = 5.0 + 0.2 * r.astype(np.float64)
    unw = ramp.astype(np.float32)
    m = np.ones((200, 10), np.uint8)
    m[90:110, :] = 0
    u0, _, _ = _bridge(unw, m, bridge_min_num_pixel=1, bridge_erosion_size=0, bridge_radius=15)
    u1, _, _ = _bridge(unw, m, bridge_min_num_pixel=1, bridge_erosion_size=0, bridge_radius=15,
                       bridge_ramp_type="linear_azimuth")
    v = m == 1
    assert (u0[v] - ramp[v]).std() > 1.0
    np.testing.assert_allclose(u1[v], ramp[v], atol=1e-3)


# ── random multi-region scenes, bit-exact against the oracle ───────────────────

def _islands(seed, shape, thr=0.25, sigma=5.0):
    rng = np.random.default_rng(seed)
    field = ndimage.gaussian_filter(rng.normal(size=shape), sigma)
    isl = field > thr * field.std()
    lab, n = ndimage.label(isl, structure=np.ones((3, 3)))
    r, c = np.mgrid[0:shape[0], 0:shape[1]]
    truth = 0.5 + 0.01 * r - 0.02 * c + 0.0003 * (r - shape[0] / 2) ** 2
    jumps = rng.integers(-4, 5, n + 1).astype(np.float64) * TWO_PI
    jumps[0] = 0.0
    unw = (truth + jumps[lab] + rng.normal(0, 0.1, shape)).astype(np.float32)
    return unw, isl, n


@pytest.mark.parametrize("seed,shape", [(1, (150, 210)), (2, (257, 131)), (3, (90, 400))])
@pytest.mark.parametrize("use_mask", [False, True])
def test_bridge_random_scenes_bitexact(gpu_ready, seed, shape, use_mask):
    unw, isl, n = _islands(seed, shape)
    assert n >= 4
    if use_mask:   # masked pixels keep their (nonzero) values
        m = isl.astype(np.uint8)
        src = unw
    else:          # invalid pixels are zero, the isce3 convention
        m = None
        src = np.where(isl, unw, 0).astype(np.float32)
    for esize, minpix in ((2, 14), (0, 1), (3, 40)):
        _bridge(src, m, bridge_erosion_size=esize, bridge_min_num_pixel=minpix, bridge_radius=60)
    _bridge(src, m, bridge_max_boundary_samples=7, bridge_radius=500)


def _diagonal_lines():
    """One-pixel diagonal (rows 0-47) and anti-diagonal (rows 48-95) lines:
    8-connected but not 4-connected (7 regions against 336), each with its
    own whole-cycle offset on a background of zeros.  They cross the region
    labeller's 64 x 16 tile borders at tile corners ((15, 63) -> (16, 64),
    (15, 127) -> (16, 128), (63, 128) -> (64, 127), (63, 192) -> (64, 191)),
    across right borders only ((4, 63) -> (5, 64), (21, 191) -> (22, 192)),
    bottom borders only and a left border only ((66, 64) -> (67, 63))."""
    nrow, ncol = 96, 256
    r, c = np.mgrid[0:nrow, 0:ncol]
    truth = 0.3 + 0.004 * r - 0.002 * c
    reg = np.zeros((nrow, ncol), np.int32)
    k = 0
    for off in (48, 59, 112, 170):
        k += 1
        for rr in range(48):
            if rr + off < ncol:
                reg[rr, rr + off] = k
    for off in (130, 191, 255):
        k += 1
        for rr in range(48, 96):
            if 0 <= off - rr < ncol:
                reg[rr, off - rr] = k
    jumps = np.array([0, 3, -2, 1, 4, -1, 2, -3], np.float64) * TWO_PI
    return np.where(reg > 0, truth + jumps[reg], 0.0).astype(np.float32), reg, truth


def test_bridge_regions_diagonal_across_tile_borders(gpu_ready):
    unw, reg, truth = _diagonal_lines()
    assert ndimage.label(reg > 0, structure=np.ones((3, 3)))[1] == 7
    u, cc, n = _bridge(unw, None, bridge_erosion_size=0, bridge_min_num_pixel=14, bridge_radius=60)
    assert n == 7
    # regions are numbered by their first raster pixel, the order the lines were drawn in
    np.testing.assert_array_equal(cc, reg.astype(cc.dtype))
    # every line is moved onto the cycle of region 1 (the largest, lowest label on a tie: +3 cycles)
    v = reg > 0
    np.testing.assert_allclose(u[v] - truth[v], 3 * TWO_PI, atol=1e-3)


@pytest.mark.parametrize("use_mask", [False, True])
def test_bridge_regions_sparse_random(gpu_ready, use_mask):
    """Random pixels at density 0.45, just above the 8-neighbour site
    percolation threshold: 216 regions joined mostly through diagonal
    links, 14 of at least 14 pixels."""
    rng = np.random.default_rng(11)
    valid = rng.random((80, 300)) < 0.45
    lab, nlab = ndimage.label(valid, structure=np.ones((3, 3)))
    r, c = np.mgrid[0:80, 0:300]
    jumps = rng.integers(-3, 4, nlab + 1) * TWO_PI
    val = (0.2 + 0.003 * r + 0.001 * c + jumps[lab]).astype(np.float32)
    if use_mask:
        _bridge(val, valid.astype(np.uint8), bridge_erosion_size=0, bridge_min_num_pixel=14, bridge_radius=60)
    else:
        _bridge(np.where(valid, val, 0).astype(np.float32), None, bridge_erosion_size=0, bridge_min_num_pixel=14,
                bridge_radius=60)


def test_bridge_regions_at_scale_known_answer(gpu_ready):
    """4096 x 4096: five bands separated by one-pixel zero rows, each crossed
    by diagonal zero lines (c - r = 301 k + 37 - 4096, c + r = 419 k + 53).
    Those lines cut a band into 415 pieces under 4-connectivity but not
    under 8-connectivity, so the bands stay whole only if every diagonal link
    across the labeller's 64 x 16 tile borders is found (about 16,000 tiles,
    one global forest).  Answer by construction: labels = band numbers (in
    raster order), and every band moved onto band 1's cycle (the largest),
    each pixel gaining float32(-J 2 pi) for its band's jump J."""
    n = 4096
    heights, jumps = [1100, 700, 900, 800, 592], [0, 3, -2, 5, -1]
    band = np.zeros((n, n), np.int32)
    r0 = 0
    for k, h in enumerate(heights):
        band[r0:r0 + h] = k + 1
        r0 += h + 1
    r, c = np.mgrid[0:n, 0:n]
    band[(np.mod(c - r + n - 37, 301) == 0) | (np.mod(c + r - 53, 419) == 0)] = 0
    J = np.array([0] + jumps, np.float64)
    unw = np.where(band > 0, 0.5 + 1e-4 * r + 2e-4 * c + J[band] * TWO_PI, 0.0).astype(np.float32)
    del r, c
    u, cc = _cuphu_ext._bridge_apply_test(unw, mask=None, bridge_radius=500, bridge_min_num_pixel=14,
                                          bridge_erosion_size=2, bridge_max_boundary_samples=4096,
                                          bridge_ramp_type=None, bridge_ramp_max_num_sample=1000000)
    np.testing.assert_array_equal(cc, band.astype(cc.dtype))
    shift = np.array([np.float32(-j * TWO_PI) for j in J], np.float32)
    np.testing.assert_array_equal(u, np.where(band > 0, unw + shift[band], unw).astype(np.float32))


@pytest.mark.parametrize("ramp", RAMPS)
def test_bridge_ramp_types_bitexact(gpu_ready, ramp):
    unw, isl, _ = _islands(7, (180, 240))
    src = np.where(isl, unw, 0).astype(np.float32)
    _bridge(src, None, bridge_ramp_type=ramp, bridge_radius=40)
    _bridge(src, None, bridge_ramp_type=ramp, bridge_radius=40, bridge_ramp_max_num_sample=500)


def test_bridge_device_entry_point(gpu_ready):
    unw, isl, _ = _islands(9, (128, 160))
    m = isl.astype(np.uint8)
    eu, ecc, _ = O.bridge(unw, m, radius=50, min_num_pixel=5, erosion_size=1)
    bp = _capi.CuPhuBridgeParams()
    _capi.lib.cuphu_default_bridge_params(C.byref(bp))
    bp.enabled, bp.radius, bp.min_num_pixel, bp.erosion_size = 1, 50, 5, 1
    du, dm = G.dev(unw), G.dev(m, np.uint8)
    dc = G.dev(np.zeros(unw.shape, np.uint32).view(np.int32))
    _capi.check(_capi.lib.cuphu_bridge_gpu(G.p(du), G.p(dc), G.p(dm), 128, 160, C.byref(bp), G.stream()), "bridge")
    np.testing.assert_array_equal(G.host(du), eu)
    np.testing.assert_array_equal(G.host(dc).view(np.uint32), ecc)


# ── unwrap() switches ─────────────────────────────────────────────────────────

def test_unwrap_bridge_switch(gpu_ready):
    s = synth.make_scene(120, 150, nbumps=2, npatches=0, nlooks=10.0, seed=41)
    mask = np.ones(s.corr.shape, np.uint8)
    mask[:, 70:76] = 0
    mask[50:56, :] = 0
    kw = dict(init="laplace", ntiles=(1, 1), reference_pixel=None, mask=mask,
              laplace_options=dict(n_irls=2, max_iter=300))
    base, cc0 = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", **kw)
    off, cc1 = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", bridge=False, **kw)
    np.testing.assert_array_equal(base, off)
    np.testing.assert_array_equal(cc0, cc1)
    br, ccb = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", bridge=True, bridge_radius=40, bridge_min_num_pixel=5,
                           **kw)
    eu, ecc, nreg = O.bridge(base, mask, radius=40, min_num_pixel=5, erosion_size=2)
    assert nreg == 4
    np.testing.assert_array_equal(br, eu)
    np.testing.assert_array_equal(ccb, ecc)
    bn, _ = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", bridge=True, bridge_radius=40, bridge_min_num_pixel=5,
                         bridge_ramp_type=None, **kw)
    np.testing.assert_array_equal(bn, br)

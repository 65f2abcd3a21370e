"""Post-passes on the GPU: isolated whole-cycle spike correction and the
Laplace tile-seam neighbour feedback (reference cuphu_solver.cpp:1385-1657).

The reference's own hook tests (test_fix_cycle_spikes.py,
test_laplace_neighbor_feedback.py) run unchanged through the
reference-named bindings, every output is compared bit for bit with the
oracle, and the API-level switches of unwrap() are checked against the
oracle applied to the unswitched result."""

import ctypes as C

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi, synth
from cuphu_amd._ext import _cuphu_ext
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWO_PI = 2.0 * np.pi


def _ramp(nrow, ncol, a, b):
    r, c = np.mgrid[0:nrow, 0:ncol]
    return a * r + b * c


def _spikes(unw, mask=None):
    out = _cuphu_ext._fix_cycle_spikes_test(unw, mask=mask)
    np.testing.assert_array_equal(out, O.fix_cycle_spikes(unw, mask))
    return out


def _feedback(unw, mask, ntr, ntc, rovl, covl, feather):
    out = _cuphu_ext._laplace_neighbor_feedback_test(unw, mask, ntilerow=ntr, ntilecol=ntc, row_ovrlp=rovl,
                                                     col_ovrlp=covl, feather_px=feather)
    np.testing.assert_array_equal(out, O.neighbor_feedback(unw, mask, ntr, ntc, rovl, covl, feather))
    return out


# ── the reference's hook tests ────────────────────────────────────────────────

def test_fix_cycle_spikes_reference_cases(gpu_ready):
    ramp = _ramp(60, 40, 0.02, 0.01)
    u = ramp.astype(np.float32).copy()
    u[27, :] -= np.float32(TWO_PI)
    np.testing.assert_allclose(_spikes(u), ramp, atol=1e-3)
    ramp = _ramp(40, 60, 0.015, 0.02)
    u = ramp.astype(np.float32).copy()
    u[:, 33] += np.float32(2.0 * TWO_PI)
    np.testing.assert_allclose(_spikes(u), ramp, atol=1e-3)
    u = _ramp(50, 50, 0.05, 0.03).astype(np.float32)
    np.testing.assert_array_equal(_spikes(u), u)
    ramp = _ramp(60, 40, 0.02, 0.01)
    u = ramp.astype(np.float32).copy()
    u[20:40, :] -= np.float32(TWO_PI)
    np.testing.assert_array_equal(_spikes(u), u)
    u = ramp.astype(np.float32).copy()
    u[27, :] -= np.float32(TWO_PI)
    m = np.ones((60, 40), np.uint8)
    m[27, :10] = 0
    out = _spikes(u, m)
    np.testing.assert_array_equal(out[27, :10], u[27, :10])
    np.testing.assert_allclose(out[27, 10:], ramp[27, 10:], atol=1e-3)


def test_neighbor_feedback_reference_cases(gpu_ready):
    unw = np.zeros((50, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    fixed = _feedback(unw, np.ones((50, 200), np.uint8), 1, 2, 0, 64, 50)
    assert abs(fixed[0, 100] - fixed[0, 99]) < 0.5
    np.testing.assert_allclose(fixed[:, :90], unw[:, :90])
    np.testing.assert_allclose(fixed[:, 160:], unw[:, 160:])
    r, c = np.mgrid[0:500, 0:200]
    unw = (0.15 * c).astype(np.float32)
    fixed = _feedback(unw, np.ones((500, 200), np.uint8), 1, 2, 0, 64, 200)
    assert np.abs(fixed[:, 100] - fixed[:, 99]).mean() <= np.abs(unw[:, 100] - unw[:, 99]).mean() * 1.1
    unw = np.zeros((60, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    m = np.ones((60, 200), np.uint8)
    m[:, 99:101] = 0
    m[::10, 99:101] = 1
    corr = _feedback(unw, m, 1, 2, 0, 64, 50)[:, 110] - unw[:, 110]
    sparse = np.delete(corr, np.arange(0, 60, 10))
    assert np.abs(sparse).min() > 1.0
    np.testing.assert_allclose(sparse, np.median(corr[::10]), atol=1.0)
    assert np.abs(np.diff(corr)).max() < 2.0
    unw = np.zeros((400, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    m = np.zeros((400, 200), np.uint8)
    m[:, 150] = 1
    m[190:200, 99:101] = 1
    corr = _feedback(unw, m, 1, 2, 0, 64, 50)[:, 150] - unw[:, 150]
    for i in (0, -1, 50, 350):
        assert abs(corr[i]) < 0.5
    unw = np.full((40, 40), 3.5, np.float32)
    np.testing.assert_array_equal(_feedback(unw, None, 1, 1, 0, 0, 200), unw)


# ── bit-exact against the oracle on random inputs ─────────────────────────────

@pytest.mark.parametrize("shape", [(3, 3), (23, 300), (301, 257), (700, 90)])
@pytest.mark.parametrize("masked", [False, True])
def test_fix_cycle_spikes_random_bitexact(gpu_ready, shape, masked):
    rng = np.random.default_rng(shape[0] * 7 + shape[1] + masked)
    nrow, ncol = shape
    u = (_ramp(nrow, ncol, 0.01, -0.02) + rng.normal(0, 0.2, shape)).astype(np.float32)
    for r in rng.choice(nrow, size=min(8, nrow), replace=False):
        u[r, :] += np.float32(TWO_PI * rng.integers(-3, 4))
    for c in rng.choice(ncol, size=min(8, ncol), replace=False):
        u[:, c] -= np.float32(TWO_PI * rng.integers(-2, 3))
    u[rng.random(shape) < 0.01] = 0.0
    u[0, 0] = -0.0
    m = (rng.random(shape) > 0.3).astype(np.uint8) if masked else None
    _spikes(u, m)


@pytest.mark.parametrize("ntr,ntc,rovl,covl,feather", [(2, 3, 40, 64, 50), (3, 2, 0, 32, 200), (4, 4, 64, 64, 30)])
def test_neighbor_feedback_random_bitexact(gpu_ready, ntr, ntc, rovl, covl, feather):
    rng = np.random.default_rng(ntr * 100 + ntc * 10 + feather)
    nrow, ncol = 330, 410
    u = (_ramp(nrow, ncol, 0.03, 0.02) + rng.normal(0, 0.3, (nrow, ncol))).astype(np.float32)
    m = (rng.random((nrow, ncol)) > 0.2).astype(np.uint8)
    m[100:250, :] = 0   # a long masked band: gaps wider than the fill reach
    m[150:160, ::7] = 1
    _feedback(u, m, ntr, ntc, rovl, covl, feather)
    _feedback(u, None, ntr, ntc, rovl, covl, feather)


def test_post_passes_device_entry_points(gpu_ready):
    """The device-pointer entry points equal the host hooks."""
    rng = np.random.default_rng(3)
    u = (_ramp(200, 180, 0.02, 0.01) + rng.normal(0, 0.1, (200, 180))).astype(np.float32)
    u[77, :] += np.float32(TWO_PI)
    m = (rng.random(u.shape) > 0.1).astype(np.uint8)
    du, dm = G.dev(u), G.dev(m, np.uint8)
    _capi.check(_capi.lib.cuphu_fix_cycle_spikes_gpu(G.p(du), G.p(dm), 200, 180, G.stream()), "spikes")
    np.testing.assert_array_equal(G.host(du), O.fix_cycle_spikes(u, m))
    du = G.dev(u)
    _capi.check(_capi.lib.cuphu_neighbor_feedback_gpu(G.p(du), G.p(dm), 200, 180, 2, 2, 30, 30, 40, G.stream()),
                "feedback")
    np.testing.assert_array_equal(G.host(du), O.neighbor_feedback(u, m, 2, 2, 30, 30, 40))


# ── unwrap() switches ─────────────────────────────────────────────────────────

def test_unwrap_post_pass_switches(gpu_ready):
    s = synth.make_scene(256, 300, nbumps=3, npatches=1, nlooks=10.0, seed=31, patch=40)
    kw = dict(init="laplace", ntiles=(2, 2), tile_overlap=48, reference_pixel=None,
              laplace_options=dict(n_irls=2, max_iter=200))
    base, cc0 = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", **kw)
    off, cc1 = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", fix_cycle_spikes=False,
                            laplace_neighbor_feedback=False, **kw)
    np.testing.assert_array_equal(base, off)
    np.testing.assert_array_equal(cc0, cc1)
    fb, _ = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", laplace_neighbor_feedback=True,
                         laplace_neighbor_feedback_feather=60, **kw)
    np.testing.assert_array_equal(fb, O.neighbor_feedback(base, None, 2, 2, 48, 48, 60))
    sp, _ = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", fix_cycle_spikes=True, laplace_neighbor_feedback=True,
                         laplace_neighbor_feedback_feather=60, **kw)
    np.testing.assert_array_equal(sp, O.fix_cycle_spikes(fb))


def test_unwrap_feedback_with_padded_mask_uses_original_mask(gpu_ready):
    """mask_pad_distance > 0: the solve uses the dilated mask without
    feedback; feedback then runs on the result with the original mask."""
    s = synth.make_scene(200, 260, nbumps=2, npatches=0, nlooks=10.0, seed=32)
    mask = np.ones(s.corr.shape, np.uint8)
    mask[90:110, 40:220] = 0
    kw = dict(init="laplace", ntiles=(2, 2), tile_overlap=40, reference_pixel=None, mask=mask, mask_pad_distance=3,
              laplace_options=dict(n_irls=1, max_iter=200))
    base, _ = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", **kw)
    fb, _ = cuphu.unwrap(s.igram, s.corr, 10.0, "defo", laplace_neighbor_feedback=True,
                         laplace_neighbor_feedback_feather=50, **kw)
    np.testing.assert_array_equal(fb, O.neighbor_feedback(base, mask, 2, 2, 40, 40, 50))

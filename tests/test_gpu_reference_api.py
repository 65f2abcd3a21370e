"""The reference's API-level tests (test_unwrap.py, test_reference_pixel.py,
test_mask_pad_distance.py) that exercise the GPU path, with the same inputs
and thresholds.  Where the reference calls the default init='mcf', which
needs SNAPHU's CPU solver, init='laplace' is used; the MCF reference-pixel
case registers a zero-flow solver on a residue-free ramp, whose optimum
flows are all zero."""

import numpy as np
import pytest

import cuphu_amd as cuphu

pytestmark = pytest.mark.gpu


@pytest.fixture()
def small_igram():
    """50x60 linear phase ramp (test_unwrap.py:24-33)."""
    r = np.arange(50, dtype=np.float32)
    c = np.arange(60, dtype=np.float32)
    return np.exp(1j * (0.1 * r[:, None] + 0.2 * c[None, :])).astype(np.complex64)


@pytest.fixture()
def small_corr():
    return np.full((50, 60), 0.9, dtype=np.float32)


def _ramp_scene():
    unw_true = 0.05 * np.arange(90)[None, :] * np.ones((60, 1)) + 10.0
    return np.exp(1j * unw_true).astype(np.complex64), np.full((60, 90), 0.8, dtype=np.float32)


def _wrapped(igram):
    phi = np.angle(igram)
    return np.where(phi < 0, phi + 2 * np.pi, phi)


def test_unwrap_phase_is_consistent(gpu_ready, small_igram, small_corr):
    """test_unwrap.py:115-126: rewrapping reproduces the input phase."""
    unw, _ = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace")
    diff = np.abs(np.angle(np.exp(1j * unw)) - np.angle(small_igram))
    diff = np.where(diff > np.pi, 2 * np.pi - diff, diff)
    assert float(np.nanpercentile(diff, 99)) < 0.1


def test_custom_reference_pixel(gpu_ready):
    """test_reference_pixel.py:74-85."""
    igram, corr = _ramp_scene()
    unw, _ = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", reference_pixel=(30, 45))
    assert unw[30, 45] == pytest.approx(_wrapped(igram)[30, 45], abs=1e-4)


def test_reference_pixel_safe_when_masked_invalid(gpu_ready):
    """test_reference_pixel.py:88-102."""
    igram, corr = _ramp_scene()
    mask = np.ones((60, 90), np.uint8)
    mask[0, 0] = 0
    unw, cc = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", mask=mask)
    assert np.isfinite(unw).all()
    assert cc[0, 0] == 0


def test_mcf_reference_pixel_is_a_noop(gpu_ready):
    """test_reference_pixel.py:59-71: an MCF result is anchored at (0, 0)
    already, so the default shift changes nothing."""
    igram, corr = _ramp_scene()
    cuphu.set_flow_solver(lambda costs, phase, init: np.zeros(costs.shape[0], np.int16))
    try:
        a, _ = cuphu.unwrap(igram, corr, nlooks=10.0, init="mcf")
        b, _ = cuphu.unwrap(igram, corr, nlooks=10.0, init="mcf", reference_pixel=None)
    finally:
        cuphu.set_flow_solver(None)
    np.testing.assert_array_equal(a, b)


def test_default_zero_matches_no_padding(gpu_ready):
    """test_mask_pad_distance.py:56-68."""
    unw_true = 0.04 * np.arange(60)[None, :] * np.ones((60, 1))
    igram = np.exp(1j * unw_true).astype(np.complex64)
    corr = np.full((60, 60), 0.8, dtype=np.float32)
    mask = np.ones((60, 60), np.uint8)
    mask[:, 25:35] = 0
    ua, ca = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", mask=mask)
    ub, cb = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", mask=mask, mask_pad_distance=0)
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(ca, cb)


def test_padding_helps_isolated_feature(gpu_ready):
    """test_mask_pad_distance.py:121-148: the padded run reports no
    validity outside the original mask."""
    r, c = np.mgrid[0:150, 0:150]
    igram = np.exp(1j * (0.03 * c + 0.02 * r).astype(np.float32)).astype(np.complex64)
    corr = np.full((150, 150), 0.85, dtype=np.float32)
    mask = np.zeros((150, 150), np.uint8)
    mask[:, :60] = 1
    mask[60:90, 60:75] = 1
    mask[60:90, 75:100] = 1
    u0, _ = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", mask=mask, mask_pad_distance=0)
    u20, cc20 = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace", mask=mask, mask_pad_distance=20)
    assert u0.shape == u20.shape == (150, 150)
    assert (cc20[mask == 0] == 0).all()


def test_gpu_info(gpu_ready):
    """test_unwrap.py:362-367."""
    info = cuphu.gpu_info(0)
    assert "name" in info and "total_memory" in info and info["total_memory"] > 0
    assert cuphu.gpu_count() >= 1 and isinstance(cuphu.gpu_name(0), str)


def test_unwrap_with_mask(gpu_ready, small_igram, small_corr):
    """test_unwrap.py:129-138: masked pixels get conncomp label 0."""
    mask = np.ones((50, 60), dtype=np.uint8)
    mask[20:30, 20:40] = 0
    unw, cc = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", mask=mask)
    assert unw.shape == (50, 60)
    assert (cc[20:30, 20:40] == 0).all()
    assert (cc[mask != 0] == 1).all()


def test_unwrap_cost_modes(gpu_ready, small_igram, small_corr):
    """test_unwrap.py:141-147 with init='laplace'; 'topo' runs the defo
    kernels in the reference (cuphu_solver.cpp:441-450), so its result equals
    'defo' bit for bit."""
    out = {}
    for mode in ("smooth", "defo", "topo"):
        out[mode], _ = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, cost=mode, init="laplace")
        assert np.isfinite(out[mode]).all()
    np.testing.assert_array_equal(out["topo"], out["defo"])


def test_reopt_single_tile_is_noop(gpu_ready, small_igram, small_corr):
    """test_unwrap.py:243-254 with init='laplace': single_tile_reoptimize has
    no effect on a (1, 1) tiling."""
    u_off, c_off = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", ntiles=(1, 1),
                                single_tile_reoptimize=False)
    u_on, c_on = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", ntiles=(1, 1),
                              single_tile_reoptimize=True)
    np.testing.assert_array_equal(u_off, u_on)
    np.testing.assert_array_equal(c_off, c_on)


def test_reopt_default_off_unchanged(gpu_ready, small_igram, small_corr):
    """test_unwrap.py:274-286 with init='laplace': omitting
    single_tile_reoptimize equals passing False."""
    u_d, c_d = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", ntiles=(2, 2), tile_overlap=4)
    u_e, c_e = cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", ntiles=(2, 2), tile_overlap=4,
                            single_tile_reoptimize=False)
    np.testing.assert_array_equal(u_d, u_e)
    np.testing.assert_array_equal(c_d, c_e)


def test_reopt_tiled_is_unsupported(gpu_ready, small_igram, small_corr):
    """Tiled single_tile_reoptimize is a full-scene CPU TreeSolve in the
    reference (SNAPHU), which this build does not have: it raises instead
    of returning an unrefined result."""
    with pytest.raises(ValueError, match="single_tile_reoptimize"):
        cuphu.unwrap(small_igram, small_corr, nlooks=10.0, init="laplace", ntiles=(2, 2), tile_overlap=4,
                     single_tile_reoptimize=True)

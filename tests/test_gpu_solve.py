"""Solver-level parity (GPU): PCG/IRLS against the oracle, the reference's
unwrap() semantics with init='laplace', batching and tiling properties.

Tolerances: the plain and batched solvers run the oracle's FP32 operation
sequence and differ only in FP64 reduction order; the fused solver takes
beta from the expansion of r.z, equal to the oracle's up to FP rounding.
All are compared to the FP64 oracle by the 2-pi ambiguity count
k = rint((u - phi)/2pi): >= 99.9 % of pixels must agree, and the FP64 modes
must agree with the FP64 oracle to 1e-5 rad on converged solves.
"""

import ctypes as C

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi, synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def amb(u, phi):
    return np.rint((u.astype(np.float64) - phi.astype(np.float64)) / TWOPI)


def _oracle_solve(s, f64=True, mask=None, **opts):
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    sig = G.sigsq_from_mask(ph, s.corr, s.nlooks, mask)[:, 1]
    fn = O.pcg_irls_f64 if f64 else O.pcg_irls_f32
    u, st = fn(sig, ph, **opts)
    return u, st, ph


SMALL = [synth.make_scene(70, 90, nbumps=2, npatches=1, nlooks=10.0, seed=21, patch=15),
         synth.make_scene(96, 64, nbumps=3, npatches=0, nlooks=5.0, seed=22)]


@pytest.mark.parametrize("k", [0, 1])
def test_plain_solver_matches_fp32_oracle(gpu_ready, k):
    s = SMALL[k]
    opts = dict(n_irls=3, max_iter=300, tol=1e-4)
    o = _capi.default_options(solver="plain", **opts)
    u, cc, st = G.batch_solve([s], o)
    uo, sto, ph = _oracle_solve(s, f64=False, **opts)
    agree = (amb(u[0], ph) == amb(uo, ph)).mean()
    print("plain vs oracle32 agreement", agree, st[0]["iters"], sto["iters"])
    assert agree >= 0.999
    assert np.abs(u[0] - uo).max() < 1e-2


@pytest.mark.parametrize("solver", ["plain", "batched", "fused"])
def test_fp64_modes_match_fp64_oracle(gpu_ready, solver):
    s = SMALL[0]
    opts = dict(n_irls=2, max_iter=4000, tol=1e-12)
    o = _capi.default_options(solver=solver, precision=64, **opts)
    u, _, st = G.batch_solve([s], o)
    uo, sto, ph = _oracle_solve(s, f64=True, **opts)
    assert all(x == 0 for x in st[0]["pass_status"]), st[0]
    np.testing.assert_allclose(u[0], uo, atol=1e-5)


@pytest.mark.parametrize("solver", ["batched", "fused"])
@pytest.mark.parametrize("k", [0, 1])
def test_batched_fp32_agrees_with_fp64_oracle(gpu_ready, k, solver):
    s = SMALL[k]
    opts = dict(n_irls=5, max_iter=1000, tol=1e-6)
    u, cc, st = G.batch_solve([s], _capi.default_options(solver=solver, **opts))
    uo, sto, ph = _oracle_solve(s, f64=True, **opts)
    agree = (amb(u[0], ph) == amb(uo, ph)).mean()
    print("batched vs oracle64", agree, st[0]["iters"], sto["iters"], np.abs(u[0] - uo).max())
    assert agree >= 0.999


@pytest.mark.parametrize("solver", ["batched", "fused"])
def test_capped_noisy_solve_tracks_fp64_with_residual_monitor(gpu_ready, solver):
    """Noisy low-looks instance stopped by the iteration cap, true residual
    recomputed every 10 iterations: the FP32 solve must keep tracking the FP64
    oracle (no replacement of an unconverged residual)."""
    s = synth.make_scene(160, 224, nbumps=3, npatches=1, nlooks=1.0, seed=77, patch=40)
    opts = dict(n_irls=3, max_iter=150, tol=1e-9)
    u, cc, st = G.batch_solve([s], _capi.default_options(solver=solver, rr_interval=10, **opts))
    assert all(x == 1 for x in st[0]["pass_status"][:4]), st[0]   # capped, not converged
    uo, sto, ph = _oracle_solve(s, f64=True, **opts)
    agree = (amb(u[0], ph) == amb(uo, ph)).mean()
    print(solver, "capped noisy vs oracle64", agree)
    assert agree >= 0.999


@pytest.mark.parametrize("solver", ["batched", "fused"])
def test_batch_equals_solo_bitwise(gpu_ready, solver):
    scenes = [synth.make_scene(40 + 13 * i, 50 + 29 * i, nbumps=2, npatches=0, nlooks=[1, 5, 10, 20][i % 4],
                               seed=100 + i) for i in range(5)]
    o = _capi.default_options(n_irls=3, max_iter=200, tol=1e-5, solver=solver)
    ub, cb, sb = G.batch_solve(scenes, o)
    for i, s in enumerate(scenes):
        us, cs, ss = G.batch_solve([s], o)
        np.testing.assert_array_equal(ub[i], us[0])
        np.testing.assert_array_equal(cb[i], cs[0])
        assert sb[i]["iters"] == ss[0]["iters"]
    og = _capi.default_options(n_irls=3, max_iter=200, tol=1e-5, group_bytes_mb=1, solver=solver)
    ug, cg, sg = G.batch_solve(scenes, og)
    for i in range(len(scenes)):
        np.testing.assert_array_equal(ug[i], ub[i])


@pytest.mark.parametrize("solver", ["batched", "fused"])
def test_graph_and_eager_launches_identical(gpu_ready, solver):
    s = SMALL[1]
    u1, c1, s1 = G.batch_solve([s], _capi.default_options(n_irls=2, max_iter=150, use_graph=1, solver=solver))
    u2, c2, s2 = G.batch_solve([s], _capi.default_options(n_irls=2, max_iter=150, use_graph=0, solver=solver))
    u3, c3, s3 = G.batch_solve([s], _capi.default_options(n_irls=2, max_iter=150, use_graph=1, solver=solver))
    np.testing.assert_array_equal(u1[0], u2[0])
    np.testing.assert_array_equal(u1[0], u3[0])
    assert s1[0]["iters"] == s2[0]["iters"] == s3[0]["iters"]


def _check_choice(tuned):
    """One candidate is chosen; a candidate other than the default shape only
    when it is at least 3 % faster than the default."""
    ms = {ln: float(ln.split(": ")[1].split(" ms")[0]) for ln in tuned}
    chosen = [ln for ln in tuned if ln.endswith("(chosen)")]
    assert len(chosen) == 1, tuned
    assert ms[chosen[0]] == min(ms.values()) or all(not ln.endswith("(default)") for ln in tuned), tuned
    default = [ln for ln in tuned if ln.endswith("(default)")]
    if default:      # the default was timed but lost
        assert ms[chosen[0]] <= 0.97 * ms[default[0]], tuned


def test_autotuned_tile_shape_agrees_with_default(gpu_ready, capfd):
    scenes = [synth.make_scene(300, 700, nbumps=2, npatches=1, nlooks=10.0, seed=31, patch=40),
              synth.make_scene(90, 200, nbumps=1, npatches=0, nlooks=5.0, seed=32)]
    opts = dict(n_irls=2, max_iter=200, tol=1e-6)
    u0, c0, s0 = G.batch_solve(scenes, _capi.default_options(**opts))
    u1, c1, s1 = G.batch_solve(scenes, _capi.default_options(autotune=1, verbose=1, **opts))
    err = capfd.readouterr().err
    tuned = [ln for ln in err.splitlines() if ln.startswith("[cuphu-autotune]")]
    # widths 700 and 200 are not multiples of 256, so only the 256-column
    # tiles of 2, 4, 8 rows per lane group are timed (a 128-column tile on
    # a wider row stride is slow), each once
    assert sorted(ln.split("rows/lane group ")[1].split(":")[0] for ln in tuned) == [
        "2, lanes/row 64", "4, lanes/row 64", "8, lanes/row 64"], err[-2000:]
    _check_choice(tuned)
    u2, c2, s2 = G.batch_solve(scenes, _capi.default_options(autotune=1, verbose=1, **opts))
    assert "[cuphu-autotune]" not in capfd.readouterr().err   # cached for the shape class
    for i, s in enumerate(scenes):
        ph = O.wrapped_phase(s.igram, correctly_rounded=True)
        assert (amb(u1[i], ph) == amb(u0[i], ph)).mean() >= 0.999
        np.testing.assert_array_equal(u2[i], u1[i])
    # widths that are multiples of 256: the 128-column tiles of 2, 4 rows are
    # candidates as well
    wide = [synth.make_scene(200, 512, nbumps=1, npatches=0, nlooks=10.0, seed=33),
            synth.make_scene(120, 256, nbumps=1, npatches=0, nlooks=5.0, seed=34)]
    G.batch_solve(wide, _capi.default_options(autotune=1, verbose=1, **opts))
    tuned = [ln for ln in capfd.readouterr().err.splitlines() if ln.startswith("[cuphu-autotune]")]
    assert sorted(ln.split("rows/lane group ")[1].split(":")[0] for ln in tuned) == [
        "2, lanes/row 32", "2, lanes/row 64", "4, lanes/row 32", "4, lanes/row 64", "8, lanes/row 64"]
    _check_choice(tuned)


def test_conncomp_of_batched_output_matches_oracle(gpu_ready):
    s = SMALL[0]
    mask = np.ones(s.corr.shape, np.uint8)
    mask[10:20, 30:45] = 0
    u, cc, _ = G.batch_solve([s], _capi.default_options(n_irls=2, max_iter=200), masks=[mask])
    prm = _capi.default_params()
    np.testing.assert_array_equal(cc[0], O.conncomp(u[0], s.corr, mask, prm.minconncompfrac, prm.maxncomps))


def test_all_masked_instance(gpu_ready):
    s = SMALL[1]
    mask = np.zeros(s.corr.shape, np.uint8)
    u, cc, st = G.batch_solve([s], _capi.default_options(n_irls=2), masks=[mask])
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    assert st[0]["pass_status"] == [3, 3, 3]
    exp = ph + np.float32(TWOPI) * np.rint((ph[0, 0] - ph) / np.float32(TWOPI))
    np.testing.assert_allclose(u[0], exp, atol=1e-5)
    assert (cc[0] == 0).all()


def test_phase_input_instance_equals_igram_instance(gpu_ready):
    s = SMALL[1]
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    o = _capi.default_options(n_irls=1, max_iter=100)
    u1, _, _ = G.batch_solve([s], o)
    u2, _, _ = G.batch_solve([s], o, phases=[ph])
    np.testing.assert_array_equal(u1[0], u2[0])


# ── unwrap() with init='laplace' (reference test semantics) ─────────────────
def test_ramp_256_closed_form(gpu_ready):
    s = synth.ramp_256()
    unw, cc = cuphu.unwrap(s.igram, s.corr, 1.0, init="laplace", ntiles=(1, 1))
    assert np.abs(unw - s.truth).max() < 1e-3
    assert (cc == 1).all()


def _ramp6090():
    unw_true = 0.05 * np.arange(90)[None, :] * np.ones((60, 1)) + 10.0
    return np.exp(1j * unw_true).astype(np.complex64), np.full((60, 90), 0.8, np.float32)


def _wrapped(ig):
    p = np.angle(ig)
    return np.where(p < 0, p + 2 * np.pi, p)


def test_reference_pixel_default_and_none(gpu_ready):
    ig, corr = _ramp6090()
    unw, _ = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace")
    assert unw[0, 0] == pytest.approx(_wrapped(ig)[0, 0], abs=1e-4)
    raw, _ = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", reference_pixel=None)
    n = (unw - raw)[0, 0] / (2 * np.pi)
    assert abs(n - round(n)) < 1e-3
    rp, rc = 30, 45
    u2, _ = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", reference_pixel=(rp, rc))
    assert u2[rp, rc] == pytest.approx(_wrapped(ig)[rp, rc], abs=1e-4)
    mask = np.ones((60, 90), np.uint8)
    mask[0, 0] = 0
    u3, cc3 = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", mask=mask)
    assert np.isfinite(u3).all() and cc3[0, 0] == 0


def test_outputs_shape_dtype_finite_and_prealloc(gpu_ready):
    r = np.arange(50, dtype=np.float32)
    c = np.arange(60, dtype=np.float32)
    ig = np.exp(1j * (0.1 * r[:, None] + 0.2 * c[None, :])).astype(np.complex64)
    corr = np.full((50, 60), 0.9, np.float32)
    ubuf = np.zeros((50, 60), np.float32)
    cbuf = np.zeros((50, 60), np.uint32)
    ou, oc = cuphu.unwrap(ig, corr, 10.0, init="laplace", unw=ubuf, conncomp=cbuf)
    assert ou is ubuf and oc is cbuf and np.isfinite(ubuf).all()
    assert ubuf.dtype == np.float32 and cbuf.dtype == np.uint32
    wrapped_back = np.angle(np.exp(1j * ubuf))
    d = np.abs(np.angle(np.exp(1j * (wrapped_back - np.angle(ig)))))
    assert float(np.percentile(d, 99)) < 0.1


@pytest.mark.parametrize("ntiles", [(1, 1), (2, 2), (2, 3)])
def test_tiled_matches_single(gpu_ready, ntiles):
    r = np.arange(50, dtype=np.float32)
    c = np.arange(60, dtype=np.float32)
    ig = np.exp(1j * (0.1 * r[:, None] + 0.2 * c[None, :])).astype(np.complex64)
    corr = np.full((50, 60), 0.9, np.float32)
    ref, _ = cuphu.unwrap(ig, corr, 10.0, init="laplace", ntiles=(1, 1))
    til, _ = cuphu.unwrap(ig, corr, 10.0, init="laplace", ntiles=ntiles, tile_overlap=4)
    assert float(np.percentile(np.abs(ref - til), 95)) < 0.5


def test_mask_pad_distance_restores_original_mask(gpu_ready):
    unw_true = 0.05 * np.arange(100)[None, :] * np.ones((100, 1))
    ig = np.exp(1j * unw_true).astype(np.complex64)
    corr = np.full((100, 100), 0.8, np.float32)
    mask = np.ones((100, 100), np.uint8)
    mask[:, 40:60] = 0
    for pad in [0, 5, 20, 50]:
        _, cc = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", mask=mask, mask_pad_distance=pad)
        assert (cc[:, 40:60] == 0).all() and (cc[:, :40] != 0).any()
    a = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", mask=mask)
    b = cuphu.unwrap(ig, corr, nlooks=10.0, init="laplace", mask=mask, mask_pad_distance=0)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_nan_input_does_not_crash(gpu_ready):
    ig = np.full((50, 60), np.nan + 0j, dtype=np.complex64)
    unw, _ = cuphu.unwrap(ig, np.full((50, 60), 0.9, np.float32), nlooks=10.0, init="laplace")
    assert np.isfinite(unw).all()


def test_decorrelated_strip_gives_two_components(gpu_ready):
    corr = np.full((20, 30), 0.8, np.float32)
    corr[:, 14:16] = 0.0
    unw = (0.05 * np.arange(30)[None, :] * np.ones((20, 1))).astype(np.float32)
    ig = np.exp(1j * unw).astype(np.complex64)
    _, cc = cuphu.unwrap(ig, corr, nlooks=10.0, cost="smooth", init="laplace")
    counts = np.bincount(cc.ravel())[1:]
    assert sorted(counts.tolist()) == [20 * 14, 20 * 14]


def test_scene_api_equals_batch_plus_oracle_stitch(gpu_ready):
    torch = G.torch_mod()
    s = synth.make_scene(300, 410, nbumps=3, npatches=1, nlooks=10.0, seed=77, patch=40)
    nrow, ncol = s.corr.shape
    o = _capi.default_options(n_irls=2, max_iter=150)
    ts, ov = 160, 32
    rf, rs = O.tile_axis(nrow, ts, ov)
    cf, cs = O.tile_axis(ncol, ts, ov)
    tiles = []
    for a in range(len(rf)):
        for b in range(len(cf)):
            sl = (slice(rf[a], rf[a] + rs[a]), slice(cf[b], cf[b] + cs[b]))
            tiles.append(synth.Scene(np.ascontiguousarray(s.igram[sl]), np.ascontiguousarray(s.corr[sl]),
                                     s.truth[sl], s.nlooks, 0))
    tu, tc, _ = G.batch_solve(tiles, o)
    eu, ec, _ = O.stitch(nrow, ncol, rf, rs, cf, cs, ov, ov, tu, tc, [t.corr for t in tiles], None)
    dig = G.dev(s.igram.view(np.float32).reshape(nrow, 2 * ncol))
    dc = G.dev(s.corr)
    du = G.empty((nrow, ncol), torch.float32)
    dl = G.empty((nrow, ncol), torch.int32)
    nt = C.c_int()
    prm = _capi.default_params()
    _capi.check(_capi.lib.cuphu_unwrap_scene_gpu(G.p(dig), G.p(dc), None, None, nrow, ncol, 10.0, 0, C.byref(prm),
                                                 ts, ov, 0, 0, C.byref(o), G.p(du), G.p(dl), None, C.byref(nt),
                                                 G.stream()), "scene")
    assert nt.value == len(rf) * len(cf)
    np.testing.assert_array_equal(G.host(du), eu)
    np.testing.assert_array_equal(G.host(dl).view(np.uint32), ec)


def _scene_tiles(s, ts, ov, mask=None):
    nrow, ncol = s.corr.shape
    rf, rs = O.tile_axis(nrow, ts, ov)
    cf, cs = O.tile_axis(ncol, ts, ov)
    tiles, tmask = [], []
    for a in range(len(rf)):
        for b in range(len(cf)):
            sl = (slice(rf[a], rf[a] + rs[a]), slice(cf[b], cf[b] + cs[b]))
            tiles.append(synth.Scene(np.ascontiguousarray(s.igram[sl]), np.ascontiguousarray(s.corr[sl]),
                                     s.truth[sl], s.nlooks, 0))
            tmask.append(np.ascontiguousarray(mask[sl]) if mask is not None else None)
    return (rf, rs, cf, cs), tiles, tmask


def _scene_gpu(s, ts, ov, o, mask=None, orig=None):
    torch = G.torch_mod()
    nrow, ncol = s.corr.shape
    dig = G.dev(s.igram.view(np.float32).reshape(nrow, 2 * ncol))
    dc = G.dev(s.corr)
    dm = G.dev(mask, np.uint8) if mask is not None else None
    do = G.dev(orig, np.uint8) if orig is not None else None
    du = G.empty((nrow, ncol), torch.float32)
    dl = G.empty((nrow, ncol), torch.int32)
    nt = C.c_int()
    prm = _capi.default_params()
    _capi.check(_capi.lib.cuphu_unwrap_scene_gpu(G.p(dig), G.p(dc), G.p(dm), G.p(do), nrow, ncol, s.nlooks, 0,
                                                 C.byref(prm), ts, ov, 0, 0, C.byref(o), G.p(du), G.p(dl), None,
                                                 C.byref(nt), G.stream()), "scene")
    return G.host(du), G.host(dl).view(np.uint32), nt.value


@pytest.mark.parametrize("mode", ["mask", "orig_mask"])
@pytest.mark.parametrize("integer", [0, 1])
def test_scene_gpu_registration_matches_oracle_stitch(gpu_ready, mode, integer):
    """The GPU pair registration (weighted medians of the overlap
    differences) gated by the solve mask or by orig_mask, with fractional or
    whole-cycle tile offsets, against the oracle's host stitch."""
    s = synth.make_scene(300, 410, nbumps=3, npatches=1, nlooks=10.0, seed=78, patch=40)
    nrow, ncol = s.corr.shape
    rng = np.random.default_rng(5)
    m = (rng.random((nrow, ncol)) > 0.1).astype(np.uint8)
    m[100:140, 120:260] = 0
    mask, orig = (m, None) if mode == "mask" else (None, m)
    o = _capi.default_options(n_irls=2, max_iter=150, stitch_integer=integer)
    ts, ov = 160, 32
    (rf, rs, cf, cs), tiles, tmask = _scene_tiles(s, ts, ov, mask)
    tconf = _scene_tiles(s, ts, ov, m)[2]
    tu, tc, _ = G.batch_solve(tiles, o, masks=tmask if mask is not None else None)
    eu, ec, ek = O.stitch(nrow, ncol, rf, rs, cf, cs, ov, ov, tu, tc, [t.corr for t in tiles], tconf,
                          laplace=not integer)
    if integer:
        assert np.array_equal(ek, np.rint(ek))
    gu, gl, nt = _scene_gpu(s, ts, ov, o, mask, orig)
    assert nt == len(rf) * len(cf)
    np.testing.assert_array_equal(gu, eu)
    np.testing.assert_array_equal(gl, ec)


def test_scene_merge_labels_is_scene_conncomp(gpu_ready):
    """merge_labels: the stitched scene is labelled as a whole (a component
    crossing tile seams keeps one label; the size threshold is taken over
    the scene), equal to the oracle's conncomp of the stitched phase; the
    phase itself is unchanged."""
    s = synth.make_scene(300, 410, nbumps=3, npatches=1, nlooks=10.0, seed=79, patch=40)
    ts, ov = 160, 32
    u0, l0, _ = _scene_gpu(s, ts, ov, _capi.default_options(n_irls=2, max_iter=150))
    u1, l1, _ = _scene_gpu(s, ts, ov, _capi.default_options(n_irls=2, max_iter=150, merge_labels=1))
    np.testing.assert_array_equal(u0, u1)
    prm = _capi.default_params()
    ref = O.conncomp(u1, s.corr, None, prm.minconncompfrac, prm.maxncomps)
    np.testing.assert_array_equal(l1, ref)
    assert l1.max() < l0.max()   # per-tile labels split the seam-crossing components


def test_nonfinite_inputs_scrubbed_like_the_reference(gpu_ready):
    """Non-finite interferogram elements (either part) and coherences are
    taken as 0 (_unwrap.py:515-523), now on the device: the result equals
    unwrap() of arrays scrubbed beforehand, bit for bit, tiled and with the
    reference pixel itself non-finite."""
    s = synth.make_scene(130, 170, nbumps=2, npatches=1, nlooks=10.0, seed=61, patch=30)
    rng = np.random.default_rng(4)
    ig = s.igram.copy()
    corr = s.corr.copy()
    idx = rng.choice(ig.size, 300, replace=False)
    ig.ravel()[idx[:100]] = np.complex64(complex(np.nan, 0.0))
    ig.ravel()[idx[100:150]] = np.complex64(complex(1.0, np.inf))
    corr.ravel()[idx[150:250]] = np.nan
    corr.ravel()[idx[250:]] = -np.inf
    ig[0, 0] = np.complex64(complex(np.nan, np.nan))
    clean_ig = np.where(np.isfinite(ig), ig, 0).astype(np.complex64)
    clean_corr = np.where(np.isfinite(corr), corr, 0).astype(np.float32)
    for kw in (dict(ntiles=(1, 1)), dict(ntiles=(2, 2), tile_overlap=32)):
        a, ca = cuphu.unwrap(ig, corr, 10.0, "defo", init="laplace", laplace_options=dict(n_irls=2, max_iter=200), **kw)
        b, cb = cuphu.unwrap(clean_ig, clean_corr, 10.0, "defo", init="laplace",
                             laplace_options=dict(n_irls=2, max_iter=200), **kw)
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(ca, cb)


def test_repeat_runs_are_deterministic(gpu_ready):
    s = SMALL[0]
    o = _capi.default_options(n_irls=2, max_iter=300)
    a, _, _ = G.batch_solve([s], o)
    b, _, _ = G.batch_solve([s], o)
    np.testing.assert_array_equal(a[0], b[0])


@pytest.mark.parametrize("opts", [
    dict(n_irls=3, max_iter=200, tol=1e-6),                    # every solve to the cap
    dict(n_irls=2, max_iter=37, tol=1e-6, rr_interval=10),     # odd cap, short chunks
    dict(n_irls=2, max_iter=400, tol=3e-3),                    # converges inside a chunk
    dict(n_irls=1, max_iter=200, tol=1e-6, use_graph=0),       # eager launches
])
def test_folded_reductions_bit_identical(gpu_ready, monkeypatch, opts):
    """The fused solver's iteration kernel with the previous launch's
    reduction folded in (small grids) gives the states, iteration counts and
    unwrapped phase of one reduction launch per iteration, bit for bit;
    alone and batched, with the live kernel timing on and off; in the plain
    tiling and in the merged one, forced on these small grids with
    CUPHU_LAPLACE_MERGE=2 (folded or not, the 520-column scene, 512 + 64
    work columns, runs its last 64 columns in 64 x 64 tiles and the
    90-column one its 128 work columns in 128 x 32 tiles)."""
    scenes = [synth.make_scene(70, 90, nbumps=2, npatches=1, nlooks=10.0, seed=41, patch=15),
              synth.make_scene(300, 520, nbumps=3, npatches=1, nlooks=5.0, seed=42, patch=60),
              synth.make_scene(33, 200, nbumps=1, npatches=0, nlooks=20.0, seed=43)]
    o = _capi.default_options(solver="fused", **opts)
    res = {}
    for fold, merge in (("0", "2"), ("1", "2"), ("0", "0"), ("1", "0")):
        monkeypatch.setenv("CUPHU_FUSED_FOLD", fold)
        monkeypatch.setenv("CUPHU_LAPLACE_MERGE", merge)
        key = fold + merge
        res[key] = [G.batch_solve(scenes, o)]
        res[key] += [G.batch_solve([sc], o) for sc in scenes[:2]]
        _capi.lib.cuphu_set_profiling(1)
        try:
            res[key].append(G.batch_solve(scenes[1:2], o))
        finally:
            _capi.lib.cuphu_set_profiling(0)
    for a, b in list(zip(res["02"], res["12"])) + list(zip(res["00"], res["10"])):
        for ua, ub in zip(a[0], b[0]):
            np.testing.assert_array_equal(ua, ub)
        for ca, cb in zip(a[1], b[1]):
            np.testing.assert_array_equal(ca, cb)
        assert [x["iters"] for x in a[2]] == [x["iters"] for x in b[2]]
        assert [x["relres"] for x in a[2]] == [x["relres"] for x in b[2]]


@pytest.mark.parametrize("solver", ["fused", "batched"])
def test_band_launch_order_bit_identical(gpu_ready, monkeypatch, solver):
    """Instances wider than 4096 columns are walked in bands of at most 2048
    columns;
    each tile's partial sums keep their raster slot, so the band order gives
    the raster order's results bit for bit (also batched with a narrow
    instance)."""
    scenes = [synth.make_scene(150, 5200, nbumps=2, npatches=1, nlooks=10.0, seed=51, patch=60),
              synth.make_scene(90, 300, nbumps=1, npatches=0, nlooks=5.0, seed=52)]
    o = _capi.default_options(solver=solver, n_irls=2, max_iter=150, tol=1e-6)
    res = {}
    for bands in ("0", "1"):
        monkeypatch.setenv("CUPHU_LAPLACE_BANDS", bands)
        res[bands] = G.batch_solve(scenes, o)
    for a, b in zip(res["0"][0], res["1"][0]):
        np.testing.assert_array_equal(a, b)
    assert [x["iters"] for x in res["0"][2]] == [x["iters"] for x in res["1"][2]]
    ph = O.wrapped_phase(scenes[0].igram, correctly_rounded=True)
    uo, _, _ = _oracle_solve(scenes[0], f64=True, n_irls=2, max_iter=150, tol=1e-6)
    assert (amb(res["1"][0][0], ph) == amb(uo, ph)).mean() >= 0.999

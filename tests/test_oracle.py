"""Pins the CPU oracle (CPU only).

Each oracle function is checked against something that does not restate
it: closed forms, dense numpy linear algebra, scipy, brute force, or the
reference's own known answers.  A plausible slip (dropped term, wrong sign
or index, transposed operand) fails at least one check here.
"""

import numpy as np
import pytest
from scipy import ndimage
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle as O
from cuphu_amd import synth

TWOPI = 2.0 * np.pi


def arc_list(nrow, ncol):
    """(tail, head) pixel indices of the flat arc layout: row arcs first."""
    idx = np.arange(nrow * ncol).reshape(nrow, ncol)
    t = np.concatenate([idx[:-1, :].ravel(), idx[:, :-1].ravel()])
    h = np.concatenate([idx[1:, :].ravel(), idx[:, 1:].ravel()])
    return t, h


def dense_laplacian(w, nrow, ncol, eps=1e-12):
    t, h = arc_list(nrow, ncol)
    n = nrow * ncol
    L = np.zeros((n, n))
    for a, (i, j) in enumerate(zip(t, h)):
        L[i, i] += w[a]
        L[j, j] += w[a]
        L[i, j] -= w[a]
        L[j, i] -= w[a]
    return L + eps * np.eye(n)


def wrap_pi(x):
    return x - TWOPI * np.rint(x / TWOPI)


# ── a1 wrapped phase ────────────────────────────────────────────────────────
def test_wrapped_phase_closed_form_and_range():
    th = np.linspace(-np.pi + 1e-3, np.pi, 4001)
    ig = np.exp(1j * th).astype(np.complex64)[None, :]
    phi = O.wrapped_phase(ig)
    assert phi.min() >= 0.0 and phi.max() < TWOPI
    expect = np.mod(np.angle(ig.astype(np.complex128)), TWOPI)
    d = np.abs(phi - expect)
    d = np.minimum(d, TWOPI - d)
    assert d.max() < 1e-6


def test_wrapped_phase_library_atan2f_within_one_ulp():
    rng = np.random.default_rng(0)
    ig = (rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))).astype(np.complex64)
    a = O.wrapped_phase(ig, correctly_rounded=True)
    b = O.wrapped_phase(ig, correctly_rounded=False)
    ulps = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1


def test_wrapped_phase_special_values():
    ig = np.array([[0 + 0j, -1 - 1e-30j, 1 + 0j, -1e-30 + 1j]], np.complex64)
    phi = O.wrapped_phase(ig)
    assert phi[0, 0] == 0.0
    assert phi[0, 2] == 0.0
    assert abs(phi[0, 3] - np.pi / 2) < 1e-6
    # a hair below the negative real axis wraps to (just under) pi, not -pi
    assert 0.0 <= phi[0, 1] < TWOPI


# ── a2 wrapped differences ──────────────────────────────────────────────────
def test_phase_diffs_signs_and_values():
    nrow, ncol = 9, 11
    r, c = np.mgrid[0:nrow, 0:ncol].astype(np.float64)
    u = 0.3 * c - 0.7 * r + 5.0        # row slope -0.7 rad/px, col slope 0.3
    phi = np.mod(u, TWOPI).astype(np.float32)
    rng, az = O.phase_diffs(phi)
    assert rng.shape == (nrow, ncol - 1) and az.shape == (nrow - 1, ncol)
    np.testing.assert_allclose(rng, 0.3 / TWOPI, atol=2e-6)        # right - left
    np.testing.assert_allclose(az, 0.7 / TWOPI, atol=2e-6)         # top - bottom
    big = np.mod(2.9 * c, TWOPI).astype(np.float32)               # 2.9 rad/px < pi
    rng2, _ = O.phase_diffs(big)
    np.testing.assert_allclose(rng2, 2.9 / TWOPI, atol=3e-6)
    assert rng2.min() >= -0.5 and rng2.max() < 0.5


# ── a4 boxcar ───────────────────────────────────────────────────────────────
@pytest.mark.parametrize("shape,kr,kc", [((40, 53), 3, 3), ((17, 9), 1, 4), ((30, 30), 0, 2), ((6, 5), 3, 3)])
def test_boxcar_matches_scipy_mirror(shape, kr, kc):
    rng = np.random.default_rng(1)
    src = rng.standard_normal(shape).astype(np.float32)
    got = O.boxcar2d(src, kr, kc)
    ref = ndimage.uniform_filter(src.astype(np.float64), size=(2 * kr + 1, 2 * kc + 1), mode="mirror")
    np.testing.assert_allclose(got, ref, atol=2e-6)


# ── a5 costs ────────────────────────────────────────────────────────────────
def _flat_phase(nrow, ncol):
    return np.zeros((nrow, ncol), np.float32)


def test_cost_rho_zero_and_full_coherence():
    phi = _flat_phase(8, 9)
    # nlooks = 1: threshold 1.2*(1.3+0.14) = 1.728 > any rho -> rho = 0 -> pow = 1
    c = O.costs(phi, np.full((8, 9), 0.9, np.float32), 1.0)
    assert (c[:, 1] == int((1 / 6 + 0.05) * 200 * 200 / 100)).all()   # 86
    # nlooks = 10, corr = 1: rho = 1 -> (1-rho)^p = 0 -> sigsq = 0.05*200^2/100 = 20
    c = O.costs(phi, np.ones((8, 9), np.float32), 10.0)
    assert (c[:, 1] == 20).all()


@pytest.mark.parametrize("cstd1,rho,expect", [(0.5, 0.5, 53), (1.0, 0.5, 36), (0.5, 0.75, 36)])
def test_cost_power_law_with_chosen_exponent(cstd1, rho, expect):
    # cstd2 = cstd3 = 0 makes rhopow = 2*cstd1 (1 or 2): sigsq = ((1-rho)^p/6 + 0.05)*400, truncated
    phi = _flat_phase(4, 4)
    corr = np.full((4, 4), rho, np.float32)
    c = O.costs(phi, corr, 10.0, cstd1=cstd1, cstd2=0.0, cstd3=0.0)
    p = 2 * cstd1
    assert int(((1 - rho) ** p / 6 + 0.05) * 400) == expect
    assert (c[:, 1] == expect).all()


@pytest.mark.parametrize("ncorrlooks,cstd2,cstd3,rho,expect", [
    # cstd2 alone: rhopow = cstd2 * ln n.  n = e^2 -> rhopow 2 (log10 would
    # give 0.87 -> 56, log2 2.89 -> 29).
    (float(np.exp(2.0)), 1.0, 0.0, 0.5, 36),
    # n = e^3, cstd2 = 1/3 -> rhopow 1: ((0.5)/6 + 0.05) * 400 = 53.3
    (float(np.exp(3.0)), 1.0 / 3.0, 0.0, 0.5, 53),
    # n = e, cstd2 = 2 -> rhopow 2 at rho 0.75: (0.0625/6 + 0.05) * 400 = 24.2
    (float(np.exp(1.0)), 2.0, 0.0, 0.75, 24),
    # cstd3 alone: rhopow = cstd3 * n.  n = 10, cstd3 = 0.1 -> 1 (cstd3 * ln n
    # would give 0.23 -> 76)
    (10.0, 0.0, 0.1, 0.5, 53),
    # n = 20, cstd3 = 0.1 -> rhopow 2 at rho 0.5 -> 36
    (20.0, 0.0, 0.1, 0.5, 36),
    # both: n = e^2, cstd2 = 0.5 (-> 1) + cstd3 = 1/e^2 (-> 1) = 2
    (float(np.exp(2.0)), 0.5, float(np.exp(-2.0)), 0.5, 36),
])
def test_cost_rhopow_log_and_linear_terms(ncorrlooks, cstd2, cstd3, rho, expect):
    """rhopow = 2 cstd1 + cstd2 ln(n) + cstd3 n (cuphu_cost.cu:44) with cstd1
    = 0 and the other two chosen so that rhopow is a whole number; each
    expected sigsq is ((1 - rho)^rhopow / 6 + 0.05) * 200^2 / 100 evaluated
    by hand.  The rho values stay above the threshold 1.2 (1.3/n + 0.14)."""
    assert rho >= 1.2 * (1.3 / ncorrlooks + 0.14)
    phi = _flat_phase(4, 5)
    corr = np.full((4, 5), rho, np.float32)
    c = O.costs(phi, corr, ncorrlooks, cstd1=0.0, cstd2=cstd2, cstd3=cstd3)
    assert (c[:, 1] == expect).all()
    cd = O.costs(phi, corr, ncorrlooks, cstd1=0.0, cstd2=cstd2, cstd3=cstd3, defo=True)
    assert (cd[:, 1] == expect).all()


def test_cost_default_constants_at_ten_looks():
    """The documented defaults (cstd 0.4 / 0.35 / 0.06) at n = 10 give
    rhopow = 0.8 + 0.35 ln 10 + 0.6 = 2.205904.  Evaluated by hand:
    0.4^2.205904 = 0.132487 -> (0.132487 / 6 + 0.05) * 400 = 28.83 -> 28 at
    rho = 0.6; 0.6^2.205904 = 0.324055 -> 41.60 -> 41 at rho = 0.4.  A
    base-10 logarithm would give rhopow 1.75 and 33 / 47."""
    phi = _flat_phase(3, 3)
    assert (O.costs(phi, np.full((3, 3), 0.6, np.float32), 10.0)[:, 1] == 28).all()
    assert (O.costs(phi, np.full((3, 3), 0.4, np.float32), 10.0)[:, 1] == 41).all()


def test_cost_threshold_masked_clamp_floor():
    phi = _flat_phase(5, 6)
    nl = 10.0
    thresh = 1.2 * (1.3 / nl + 0.14)
    lo = np.full((5, 6), np.float32(thresh) - np.float32(1e-3), np.float32)
    assert (O.costs(phi, lo, nl)[:, 1] == 86).all()        # below threshold -> rho = 0
    w = np.ones(O.narcs(5, 6), np.int16)
    w[3] = 0
    c = O.costs(phi, np.full((5, 6), 0.8, np.float32), nl, weights=w)
    assert tuple(c[3]) == (16000, 32000)
    cd = O.costs(phi, np.full((5, 6), 0.8, np.float32), nl, weights=w, defo=True)
    assert tuple(cd[3]) == (16000, 32000, 32000, 0)
    big = O.costs(phi, np.full((5, 6), 0.5, np.float32), nl, costscale=1e-4)
    assert (big[:, 1] == 32767).all()                       # short clamp
    tiny = O.costs(phi, np.full((5, 6), 0.5, np.float32), nl, costscale=1e6)
    assert (tiny[:, 1] == 1).all()                          # sigsqshortmin floor


def test_cost_offsets_truncate_smooth_round_defo():
    # one arc with dpsi = -0.0135 cycles and zero smoothed gradient except
    # itself: use a 1x2 image so the boxcar averages a single sample.
    nsc = 200.0
    d = -0.0135
    phi = np.array([[0.0, np.mod(d * TWOPI, TWOPI)]], np.float32)
    corr = np.full((1, 2), 0.9, np.float32)
    rng, _ = O.phase_diffs(phi)
    dp = float(rng[0, 0])
    # rho = 0.9 > threshold at nlooks 10 -> offset = nsc*(dp - avg) with avg = dp -> 0
    cs = O.costs(phi, corr, 10.0)
    assert cs[0, 0] == 0
    # rho = 0 (nlooks 1) -> offset = nsc*(dp - dp/2) = nsc*dp/2 = -1.35
    cs = O.costs(phi, corr, 1.0)
    cd = O.costs(phi, corr, 1.0, defo=True)
    v = nsc * (dp - 0.5 * dp)
    assert cs[0, 0] == int(v) == -1                          # truncation toward zero
    assert cd[0, 0] == int(np.rint(v)) == -1
    assert cd[0, 2] == 240 and cd[0, 3] == int(np.sqrt(240 * cd[0, 1]))
    phi2 = np.array([[0.0, np.mod(-0.02 * TWOPI, TWOPI)]], np.float32)
    v2 = nsc * 0.5 * float(O.phase_diffs(phi2)[0][0, 0])
    assert O.costs(phi2, corr, 1.0)[0, 0] == int(v2) == -1
    assert O.costs(phi2, corr, 1.0, defo=True)[0, 0] == int(np.rint(v2)) == -2


def test_cost_windows_are_transposed_for_azimuth():
    rng0 = np.random.default_rng(3)
    phi = rng0.uniform(0, TWOPI, (12, 15)).astype(np.float32)
    corr = np.full((12, 15), 0.9, np.float32)
    kperp, kpar = 3, 7
    c = O.costs(phi, corr, 1.0, kperp=kperp, kpar=kpar)   # rho = 0: offset = nsc*(dp - avg/2)
    r_, a_ = O.phase_diffs(phi)
    ar = ndimage.uniform_filter(r_.astype(np.float64), size=(2 * (kperp // 2) + 1, 2 * (kpar // 2) + 1), mode="mirror")
    aa = ndimage.uniform_filter(a_.astype(np.float64), size=(2 * (kpar // 2) + 1, 2 * (kperp // 2) + 1), mode="mirror")
    nra = 11 * 15
    exp_row = np.trunc(200 * (a_.astype(np.float64) - 0.5 * aa)).ravel()
    exp_col = np.trunc(200 * (r_.astype(np.float64) - 0.5 * ar)).ravel()
    assert np.abs(c[:nra, 0] - exp_row).max() <= 1
    assert np.abs(c[nra:, 0] - exp_col).max() <= 1
    assert (c[:nra, 0] == exp_row).mean() > 0.99 and (c[nra:, 0] == exp_col).mean() > 0.99


# ── a3 residues ─────────────────────────────────────────────────────────────
def test_residue_vortex_and_conservative_field():
    nrow, ncol = 24, 30
    r, c = np.mgrid[0:nrow, 0:ncol].astype(np.float64)
    vort = np.mod(np.arctan2(r - 10.5, c - 12.5), TWOPI).astype(np.float32)
    res = O.residues(vort)
    nz = np.argwhere(res != 0)
    assert nz.tolist() == [[10, 12]]
    assert abs(int(res[10, 12])) == 1
    anti = np.mod(-np.arctan2(r - 10.5, c - 12.5), TWOPI).astype(np.float32)
    assert O.residues(anti)[10, 12] == -res[10, 12]
    smooth = np.mod(0.4 * c + 0.2 * r + 0.5 * np.sin(r / 3), TWOPI).astype(np.float32)
    assert not O.residues(smooth).any()


def test_residue_equals_minus_curl_of_integer_flows():
    rng = np.random.default_rng(5)
    u = np.cumsum(np.cumsum(rng.normal(0, 1.4, (20, 25)), 0), 1) * 0.3
    phi = np.mod(u, TWOPI).astype(np.float32)
    rngd, az = O.phase_diffs(phi)
    phi64 = phi.astype(np.float64)
    # integer flows: how many cycles the true gradient differs from the wrapped one
    n_r = np.rint((np.diff(phi64, axis=1) / TWOPI) - rngd)         # (right - left) arcs
    n_a = np.rint(((phi64[:-1] - phi64[1:]) / TWOPI) - az)          # (top - bottom) arcs
    curl = (n_r[:-1, :] - n_a[:, 1:] - n_r[1:, :] + n_a[:, :-1])
    np.testing.assert_array_equal(O.residues(phi), -curl.astype(np.int8))


# ── a7/a8 assemble + matvec ─────────────────────────────────────────────────
def test_assemble_and_matvec_match_dense_laplacian():
    rng = np.random.default_rng(7)
    nrow, ncol = 7, 9
    na = O.narcs(nrow, ncol)
    w = rng.uniform(0.01, 0.2, na).astype(np.float32)
    w[::7] = 0.0
    phi = rng.uniform(0, TWOPI, (nrow, ncol)).astype(np.float32)
    x = rng.standard_normal((nrow, ncol)).astype(np.float32)
    L = dense_laplacian(w.astype(np.float64), nrow, ncol)
    y = O.matvec(w, x)
    np.testing.assert_allclose(y.ravel(), L @ x.ravel().astype(np.float64), rtol=1e-5, atol=1e-6)
    b, diag = O.assemble(w, phi)
    t, h = arc_list(nrow, ncol)
    g = wrap_pi(phi.ravel()[h].astype(np.float64) - phi.ravel()[t])
    bb = np.zeros(nrow * ncol)
    np.add.at(bb, h, w * g)
    np.add.at(bb, t, -w * g)
    np.testing.assert_allclose(b.ravel(), bb, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(diag.ravel(), np.diag(L), rtol=1e-6)
    assert abs(float(b.astype(np.float64).sum())) < 1e-4


def test_reweight_indices_and_formula():
    rng = np.random.default_rng(8)
    nrow, ncol = 6, 8
    na = O.narcs(nrow, ncol)
    sig = rng.integers(1, 200, na).astype(np.int16)
    sig[::5] = 32000
    phi = rng.uniform(0, TWOPI, (nrow, ncol)).astype(np.float32)
    u = (phi + TWOPI * rng.integers(-2, 3, (nrow, ncol))).astype(np.float32) + rng.normal(0, 0.2, (nrow, ncol)).astype(np.float32)
    w = O.reweight(sig, phi, u, 0.125)
    t, h = arc_list(nrow, ncol)
    g = wrap_pi(phi.ravel()[h].astype(np.float64) - phi.ravel()[t])
    e = (u.ravel()[h].astype(np.float64) - u.ravel()[t]) - g
    w0 = np.where(sig >= 32000, 0.0, 1.0 / sig)
    np.testing.assert_allclose(w, w0 / np.sqrt(e * e + 0.125 ** 2), rtol=2e-6)


# ── a10 PCG/IRLS ────────────────────────────────────────────────────────────
def _anchor_congruence(u, phi):
    u = u + (float(phi.ravel()[0]) - u.ravel()[0])
    return phi + TWOPI * np.rint((u - phi) / TWOPI)


def test_pcg_single_solve_matches_dense_solution():
    s = synth.make_scene(14, 17, nbumps=2, npatches=0, nlooks=5.0, seed=11)
    phi = O.wrapped_phase(s.igram)
    sig = O.costs(phi, s.corr, 5.0)[:, 1]
    u, st = O.pcg_irls_f64(sig, phi, n_irls=0, max_iter=5000, tol=1e-13)
    assert st["status"][0] == 0
    w = np.where((sig <= 0) | (sig >= 32000), 0.0, 1.0 / sig.astype(np.float64))
    L = dense_laplacian(w, 14, 17)
    t, h = arc_list(14, 17)
    g = wrap_pi(phi.ravel()[h].astype(np.float64) - phi.ravel()[t])
    b = np.zeros(14 * 17)
    np.add.at(b, h, w * g)
    np.add.at(b, t, -w * g)
    x = np.linalg.lstsq(L, b, rcond=None)[0].reshape(14, 17)
    ref = _anchor_congruence(x, phi.astype(np.float64))
    np.testing.assert_allclose(u, ref, atol=1e-6)


def _noisy_scene(seed, nrow, ncol):
    """A small scene with residues (a ramp plus 1.1 rad phase noise), so the
    IRLS reweighting changes the solution: on a residue-free field every
    arc's residual is zero and the weights only scale uniformly."""
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nrow, 0:ncol]
    truth = 0.9 * c - 0.4 * r + rng.normal(0, 1.1, (nrow, ncol))
    phi = np.mod(truth, TWOPI).astype(np.float32)
    corr = rng.uniform(0.3, 0.95, (nrow, ncol)).astype(np.float32)
    assert np.abs(O.residues(phi)).sum() >= 10
    return phi, corr


def _dense_irls(sig, phi, n_irls, dmax, dmin):
    """The IRLS of cuphu_laplace.cu:413-522 as dense float64 solves: each
    pass's weights come from the previous pass's exact solution with
    delta_k = max(dmin, dmax 0.5^k).  Returns the anchored continuous
    solution (u[0] = phi[0], cuphu_laplace.cu:534-539)."""
    nrow, ncol = phi.shape
    t, h = arc_list(nrow, ncol)
    g = wrap_pi(phi.ravel()[h] - phi.ravel()[t])
    w0 = np.where((sig <= 0) | (sig >= 32000), 0.0, 1.0 / sig.astype(np.float64))
    w = w0.copy()
    x = np.zeros(nrow * ncol)
    for outer in range(n_irls + 1):
        L = dense_laplacian(w, nrow, ncol)
        b = np.zeros(nrow * ncol)
        np.add.at(b, h, w * g)
        np.add.at(b, t, -w * g)
        x = np.linalg.solve(L, b)
        if outer < n_irls:
            d = max(dmin, dmax * 0.5 ** outer)
            e = (x[h] - x[t]) - g
            w = w0 / np.sqrt(e * e + d * d)
    x = x.reshape(nrow, ncol)
    return x + (phi[0, 0] - x[0, 0])


IRLS_CASES = [                      # the delta_min floor binds in the first and last case
    (12, (12, 13), 10.0, 4, 0.5, 0.1),
    (21, (15, 11), 5.0, 4, 0.5, 0.05),
    (33, (9, 16), 20.0, 3, 0.8, 0.3),
]


@pytest.mark.parametrize("seed,shape,nlooks,n_irls,dmax,dmin", IRLS_CASES)
def test_irls_schedule_matches_dense_irls(seed, shape, nlooks, n_irls, dmax, dmin):
    phi, corr = _noisy_scene(seed, *shape)
    sig = O.costs(phi, corr, nlooks)[:, 1]
    opts = dict(n_irls=n_irls, delta_max=dmax, delta_min=dmin, max_iter=5000, tol=1e-14)
    u, _ = O.pcg_irls_f64(sig, phi, continuous=True, **opts)
    ref = _dense_irls(sig, phi.astype(np.float64), n_irls, dmax, dmin)
    np.testing.assert_allclose(u, ref, atol=1e-7)
    uc, _ = O.pcg_irls_f64(sig, phi, **opts)
    np.testing.assert_allclose(uc, _anchor_congruence(ref, phi.astype(np.float64)), atol=1e-7)


def test_irls_reweighting_moves_the_solution():
    """The IRLS passes change the continuous solution on these scenes by far
    more than the tolerances above, so a wrong delta schedule (exponent,
    floor) or a missing reweight is visible."""
    phi, corr = _noisy_scene(12, 12, 13)
    sig = O.costs(phi, corr, 10.0)[:, 1]
    p64 = phi.astype(np.float64)
    a = _dense_irls(sig, p64, 0, 0.5, 0.1)
    b = _dense_irls(sig, p64, 3, 0.5, 0.1)
    c = _dense_irls(sig, p64, 3, 0.5, 0.2)
    d = _dense_irls(sig, p64, 3, 0.25, 0.1)
    assert np.abs(a - b).max() > 0.05
    assert np.abs(b - c).max() > 0.01 and np.abs(b - d).max() > 0.01


@pytest.mark.parametrize("seed,shape,nlooks,n_irls,dmax,dmin", IRLS_CASES)
def test_irls_f32_driver_matches_dense_irls(seed, shape, nlooks, n_irls, dmax, dmin):
    """The FP32-vector driver (o_pcg_irls_f32, cuphu_laplace.cu:413-522)
    against the dense float64 IRLS: FP32 vectors converge to ~1e-6
    relative, so the continuous solution agrees to 1e-3 rad and, after
    congruence, every 2 pi ambiguity is equal.  Each Jacobi-PCG solve
    converges within as many iterations as there are pixels, as conjugate
    gradients must: a wrong beta or step degrades it to a slower
    iteration."""
    phi, corr = _noisy_scene(seed, *shape)
    n = shape[0] * shape[1]
    sig = O.costs(phi, corr, nlooks)[:, 1]
    opts = dict(n_irls=n_irls, delta_max=dmax, delta_min=dmin, max_iter=4000, tol=1e-6)
    u, st = O.pcg_irls_f32(sig, phi, continuous=True, **opts)
    assert st["npass"] == n_irls + 1
    assert all(s_ == 0 for s_ in st["status"]), st
    assert max(st["iters"]) <= n, st
    phi64 = phi.astype(np.float64)
    ref = _dense_irls(sig, phi64, n_irls, dmax, dmin)
    assert np.abs(u - ref).max() < 1e-3
    frac = (ref - phi64) / TWOPI
    assert np.abs(np.abs(frac - np.floor(frac)) - 0.5).min() > 1e-2     # no ambiguity near a tie
    uc, _ = O.pcg_irls_f32(sig, phi, **opts)
    np.testing.assert_array_equal(np.rint((uc - phi64) / TWOPI), np.rint(frac))
    np.testing.assert_array_equal(uc, _anchor_congruence_f32(u, phi))


def _anchor_congruence_f32(u, phi):
    """FP32 congruence of an anchored field, cuphu_laplace.cu:256-266."""
    k = np.rint((u - phi) / np.float32(TWOPI)).astype(np.float32)
    return (phi + np.float32(TWOPI) * k).astype(np.float32)


def test_ramp_closed_form_both_precisions():
    """Config 1's 256^2 ramp (uniform weights, conservative gradients): the
    WLS solution is the ramp itself, unw = pi (x + y) + 6 pi, through both
    the FP64 and the FP32-vector drivers."""
    s = synth.ramp_256()
    phi = O.wrapped_phase(s.igram)
    sig = O.costs(phi, s.corr, 1.0)[:, 1]
    assert (sig == 86).all()
    amb_truth = np.rint((s.truth - phi) / TWOPI)
    u64, _ = O.pcg_irls_f64(sig, phi, n_irls=0, max_iter=3000, tol=1e-10)
    assert np.abs(u64 - s.truth).max() < 1e-5
    np.testing.assert_array_equal(np.rint((u64 - phi) / TWOPI), amb_truth)
    u32, st = O.pcg_irls_f32(sig, phi, n_irls=0, max_iter=3000, tol=1e-6)
    assert st["status"][0] in (0, 1)
    assert np.abs(u32 - s.truth).max() < 1e-4
    np.testing.assert_array_equal(np.rint((u32 - phi) / TWOPI), amb_truth)
    # with IRLS passes the ramp stays the fixed point (zero residual arcs)
    u32i, _ = O.pcg_irls_f32(sig, phi, n_irls=2, max_iter=3000, tol=1e-6)
    np.testing.assert_array_equal(np.rint((u32i - phi) / TWOPI), amb_truth)


def test_all_masked_scene_skips_solve():
    phi = np.random.default_rng(2).uniform(0, TWOPI, (6, 7)).astype(np.float32)
    sig = np.full(O.narcs(6, 7), 32000, np.int16)
    u, st = O.pcg_irls_f32(sig, phi, n_irls=2)
    assert st["status"] == [3, 3, 3] and st["iters"] == [0, 0, 0]
    exp = phi + np.float32(TWOPI) * np.rint((phi[0, 0] - phi) / np.float32(TWOPI))
    np.testing.assert_allclose(u, exp, atol=1e-5)


def test_pcg_step_reproduces_dense_pcg_iterates():
    """o_pcg_step_f32 driven 6 times against textbook Jacobi-PCG in dense
    float64 (a dropped term, sign or index in any of the updates fails)."""
    rng = np.random.default_rng(9)
    nrow, ncol = 10, 12
    na = O.narcs(nrow, ncol)
    w = rng.uniform(0.02, 0.1, na).astype(np.float32)
    phi = rng.uniform(0, TWOPI, (nrow, ncol)).astype(np.float32)
    b, diag = O.assemble(w, phi)
    dinv = (1.0 / diag).astype(np.float32)
    x = np.zeros_like(b)
    r = b.copy()
    p = np.zeros_like(b)
    gam = float((r.astype(np.float64) * (r * dinv)).sum())
    beta = 0.0
    pqs = []
    for _ in range(6):
        x, r, p, q, pq, alpha, rzn, st = O.pcg_step_f32(w, dinv, gam, beta, x, r, p)
        assert st == 0
        np.testing.assert_array_equal(q, O.matvec(w, p))
        pqs.append(pq)
        beta = rzn / gam
        gam = rzn
    L = dense_laplacian(w.astype(np.float64), nrow, ncol)
    bb = b.ravel().astype(np.float64)
    D = dinv.ravel().astype(np.float64)
    xx = np.zeros_like(bb)
    rr = bb.copy()
    zz = rr * D
    pp = zz.copy()
    rz = rr @ zz
    dpqs = []
    for _ in range(6):
        q = L @ pp
        dpqs.append(pp @ q)
        a = rz / dpqs[-1]
        xx += a * pp
        rr -= a * q
        zz = rr * D
        rzn = rr @ zz
        pp = zz + (rzn / rz) * pp
        rz = rzn
    np.testing.assert_allclose(x.ravel(), xx, rtol=2e-3, atol=2e-4 * np.abs(xx).max())
    np.testing.assert_allclose(pqs, dpqs, rtol=1e-3)
    # the step's rz is the r.z of its returned residual
    assert gam == pytest.approx(float((r.astype(np.float64) * (r * dinv)).sum()), rel=1e-12)


def test_pcg_fused_step_reproduces_dense_pcg_iterates():
    """o_pcg_fused_step_f32 driven with the fused solver's scalar recurrences
    (alpha = rz / p.q, beta from the rz expansion) against textbook
    Jacobi-PCG in dense float64; the expansion must match the r.z of the
    vectors actually formed to FP32 rounding, and q must be A p exactly."""
    rng = np.random.default_rng(19)
    nrow, ncol = 11, 13
    na = O.narcs(nrow, ncol)
    w = rng.uniform(0.02, 0.1, na).astype(np.float32)
    phi = rng.uniform(0, TWOPI, (nrow, ncol)).astype(np.float32)
    b, diag = O.assemble(w, phi)
    dinv = (1.0 / diag).astype(np.float32)
    z0 = np.zeros_like(b)
    x, r, p, q, sums = O.pcg_fused_step_f32(w, dinv, 0.0, 0.0, z0, b, z0, z0, start=True)
    np.testing.assert_array_equal(r, b)
    np.testing.assert_array_equal(q, O.matvec(w, p))
    for _ in range(6):
        pq, rz, qr, qq = sums
        alpha = rz / pq
        rz_next = rz - 2 * alpha * qr + alpha * alpha * qq
        x, r, p, q, sums = O.pcg_fused_step_f32(w, dinv, alpha, rz_next / rz, x, r, p, q)
        np.testing.assert_array_equal(q, O.matvec(w, p))
        assert sums[1] == pytest.approx(rz_next, rel=1e-5)
    L = dense_laplacian(w.astype(np.float64), nrow, ncol)
    bb = b.ravel().astype(np.float64)
    D = dinv.ravel().astype(np.float64)
    xx = np.zeros_like(bb)
    rr = bb.copy()
    zz = rr * D
    pp = zz.copy()
    rz = rr @ zz
    for _ in range(6):
        qv = L @ pp
        a = rz / (pp @ qv)
        xx += a * pp
        rr -= a * qv
        zz = rr * D
        rzn = rr @ zz
        pp = zz + (rzn / rz) * pp
        rz = rzn
    np.testing.assert_allclose(x.ravel(), xx, rtol=2e-3, atol=2e-4 * np.abs(xx).max())
    # a replacement step takes r from rt and leaves everything else as the update
    rt = (r + 0.25).astype(np.float32)
    x2, r2, p2, q2, _ = O.pcg_fused_step_f32(w, dinv, 0.5, 0.0, x, r, p, q, rt=rt)
    np.testing.assert_array_equal(r2, rt)
    np.testing.assert_array_equal(p2, rt * dinv)


def test_pcg_step_reports_breakdown():
    nrow, ncol = 3, 4
    w = np.zeros(O.narcs(nrow, ncol), np.float32)
    dinv = np.ones((nrow, ncol), np.float32)
    z = np.zeros((nrow, ncol), np.float32)
    *_, st = O.pcg_step_f32(w, dinv, 1.0, 0.0, z, z, z)
    assert st == 2   # p = r*dinv = 0 -> p.q = 0


# ── a14 conncomp ────────────────────────────────────────────────────────────
def _scipy_components(unw, corr, mask, minfrac, maxn):
    nrow, ncol = unw.shape
    valid = ~(corr < 0.05)
    if mask is not None:
        valid &= mask.astype(bool)
    h_ok = valid[:, :-1] & valid[:, 1:] & (np.abs(unw[:, :-1] - unw[:, 1:]) < np.float32(np.pi))
    v_ok = valid[:-1, :] & valid[1:, :] & (np.abs(unw[:-1, :] - unw[1:, :]) < np.float32(np.pi))
    idx = np.arange(nrow * ncol).reshape(nrow, ncol)
    ri = np.concatenate([idx[:, :-1][h_ok], idx[:-1, :][v_ok]])
    ci = np.concatenate([idx[:, 1:][h_ok], idx[1:, :][v_ok]])
    g = coo_matrix((np.ones(ri.size), (ri, ci)), shape=(nrow * ncol, nrow * ncol))
    _, lab = connected_components(g, directed=False)
    lab = lab.reshape(nrow, ncol)
    min_size = max(1, int(minfrac * nrow * ncol))
    out = np.zeros((nrow, ncol), np.uint32)
    nxt = 1
    seen = {}
    for i in range(nrow * ncol):          # components by increasing minimum index
        rr, cc = divmod(i, ncol)
        if not valid[rr, cc]:
            continue
        l = lab[rr, cc]
        if l in seen:
            continue
        members = (lab == l) & valid
        seen[l] = 0
        if members.sum() >= min_size and nxt - 1 < maxn:
            seen[l] = nxt
            nxt += 1
    for l, v in seen.items():
        out[(lab == l) & valid] = v
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_conncomp_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    unw = np.cumsum(rng.normal(0, 1.4, (23, 31)), axis=1).astype(np.float32)
    corr = rng.uniform(0, 1, (23, 31)).astype(np.float32)
    corr[rng.uniform(0, 1, corr.shape) < 0.1] = 0.01
    mask = (rng.uniform(0, 1, corr.shape) > 0.05).astype(np.uint8)
    for minfrac, maxn in [(0.0, 1000), (0.01, 32), (0.02, 3)]:
        got = O.conncomp(unw, corr, mask, minfrac, maxn)
        np.testing.assert_array_equal(got, _scipy_components(unw, corr, mask, minfrac, maxn))


def test_conncomp_reference_known_answers():
    # test_conncomp.py:32-81 fixtures
    unw = np.zeros((20, 30), np.float32)
    unw[:, 15:] += 10.0
    corr = np.full((20, 30), 0.8, np.float32)
    out = O.conncomp(unw, corr, None, 0.01, 32)
    assert sorted(set(np.unique(out)) - {0}) == [1, 2]
    assert (O.conncomp(np.zeros((10, 10), np.float32), corr[:10, :10], np.zeros((10, 10), np.uint8), 0.0, 32) == 0).all()
    u2 = np.zeros((20, 20), np.float32)
    u2[10, 10] = 100.0
    assert O.conncomp(u2, np.full((20, 20), 0.9, np.float32), None, 0.05, 32)[10, 10] == 0
    u3 = np.zeros((20, 30), np.float32)
    u3[:, 10:20] += 20.0
    u3[:, 20:] += 40.0
    o3 = O.conncomp(u3, np.full((20, 30), 0.9, np.float32), None, 0.0, 1)
    assert 1 in o3 and 2 not in o3


# ── a13 integration ─────────────────────────────────────────────────────────
def test_integrate_phase_recovers_field_from_integer_flows():
    rng = np.random.default_rng(13)
    nrow, ncol = 15, 19
    u = np.cumsum(np.cumsum(rng.normal(0, 1.6, (nrow, ncol)), 0), 1) * 0.4
    phi = np.mod(u, TWOPI).astype(np.float32)
    p64 = phi.astype(np.float64)
    wr_r = wrap_pi(np.diff(p64, axis=0))
    wr_c = wrap_pi(np.diff(p64, axis=1))
    hfl = np.rint((wr_r - np.diff(u, axis=0)) / TWOPI).astype(np.int16)   # downward subtracts
    vfl = np.rint((np.diff(u, axis=1) - wr_c) / TWOPI).astype(np.int16)
    got = O.integrate_phase(phi, hfl, vfl)
    np.testing.assert_allclose(got, u - u[0, 0] + p64[0, 0], atol=2e-4)
    zero = O.integrate_phase(phi, np.zeros_like(hfl), np.zeros_like(vfl))
    smooth = np.mod(0.3 * np.arange(ncol)[None, :] + 0.1 * np.arange(nrow)[:, None], TWOPI).astype(np.float32)
    z2 = O.integrate_phase(smooth, np.zeros_like(hfl), np.zeros_like(vfl))
    exp = 0.3 * np.arange(ncol)[None, :] + 0.1 * np.arange(nrow)[:, None]
    np.testing.assert_allclose(z2, exp, atol=2e-5)
    assert zero.shape == phi.shape


# ── a15 tiling + stitch ─────────────────────────────────────────────────────
def test_tile_axis_known_answers():
    assert O.tile_axis(8192, 2048, 256) == ([0, 1792, 3584, 5376, 7168], [2048, 2048, 2048, 2048, 1024])
    assert O.tile_axis(100, 0, 4, 3) == ([0, 32, 64], [36, 36, 36])
    assert O.tile_axis(500, 0, 0, 1) == ([0], [500])
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(50, 5000))
        nt = int(rng.integers(1, 9))
        ov = int(rng.integers(0, 40))
        f, s = O.tile_axis(n, 0, ov, nt)
        assert f[0] == 0 and f[-1] + s[-1] == n
        for k in range(nt - 1):
            assert f[k] + s[k] - f[k + 1] == ov          # exact overlap between neighbours
        ts = int(rng.integers(ov + 16, ov + 600))
        f, s = O.tile_axis(n, ts, ov)
        assert f[0] == 0 and f[-1] + s[-1] == n and all(x <= ts for x in s)
        for k in range(len(f) - 1):
            assert f[k] + s[k] - f[k + 1] == ov


def _wmedian_sort(v, w):
    order = np.argsort(v, kind="stable")
    cw = np.cumsum(w[order].astype(np.float64))
    i = int(np.searchsorted(cw, cw[-1] / 2.0, side="left"))
    return float(v[order][i])


def test_stitch_two_tiles_fractional_and_integer_offsets():
    rng = np.random.default_rng(21)
    nrow, ncol, ov = 12, 30, 6
    rf, rs = [0], [nrow]
    cf, cs = O.tile_axis(ncol, 0, ov, 2)
    base = np.cumsum(rng.normal(0, 0.3, (nrow, ncol)), axis=1).astype(np.float32)
    corr = rng.uniform(0.04, 1.0, (nrow, ncol)).astype(np.float32)
    tiles_u = [base[:, cf[0]:cf[0] + cs[0]].copy(), base[:, cf[1]:cf[1] + cs[1]] + np.float32(-7.3)]
    tiles_c = [corr[:, cf[0]:cf[0] + cs[0]].copy(), corr[:, cf[1]:cf[1] + cs[1]].copy()]
    cc = [np.ones((nrow, cs[0]), np.uint32), np.where(rng.uniform(size=(nrow, cs[1])) < 0.2, 0, 2).astype(np.uint32)]
    unw, lab, tk = O.stitch(nrow, ncol, rf, rs, cf, cs, 0, ov, tiles_u, cc, tiles_c, laplace=True)
    ovl_l = tiles_u[0][:, -ov:]
    ovl_r = tiles_u[1][:, :ov]
    ok = (tiles_c[0][:, -ov:] >= 0.05) & (tiles_c[1][:, :ov] >= 0.05)
    dv = (ovl_l - ovl_r)[ok]
    wv = np.minimum(tiles_c[0][:, -ov:], tiles_c[1][:, :ov])[ok]
    assert tk[1] == pytest.approx(_wmedian_sort(dv, wv) / TWOPI, abs=1e-12)
    assert tk[1] == pytest.approx(7.3 / TWOPI, abs=1e-5)
    np.testing.assert_allclose(unw, base, atol=1e-4)
    # labels: tile 1 offset by tile 0's max (1); zeros stay zero
    right = lab[:, cf[1] + ov // 2:]
    assert set(np.unique(right)) <= {0, 3}
    unw_i, _, tki = O.stitch(nrow, ncol, rf, rs, cf, cs, 0, ov, tiles_u, cc, tiles_c, laplace=False)
    assert tki[1] == round(7.3 / TWOPI)


def test_weighted_median_ties_and_selection():
    rng = np.random.default_rng(4)
    for trial in range(30):
        n = int(rng.integers(1, 400))
        v = rng.integers(-5, 6, n).astype(np.float32) * np.float32(0.7)
        w = rng.uniform(0.05, 1.0, n).astype(np.float32)
        nrow, ncol, ov = n, 8, 1
        u0 = np.zeros((n, 4), np.float32)
        u1 = np.zeros((n, 5), np.float32)
        u0[:, -1] = v
        c0 = np.ones((n, 4), np.float32)
        c1 = np.ones((n, 5), np.float32)
        c0[:, -1] = w
        c1[:, 0] = w
        cc = [np.zeros((n, 4), np.uint32), np.zeros((n, 5), np.uint32)]
        _, _, tk = O.stitch(nrow, ncol, [0], [n], [0, 3], [4, 5], 0, ov, [u0, u1], cc, [c0, c1])
        assert tk[1] == pytest.approx(_wmedian_sort(v, w) / TWOPI, abs=1e-12)


# ── WLS (differentiable path) ───────────────────────────────────────────────
def test_wls_solve_and_adjoint_against_dense_and_finite_differences():
    rng = np.random.default_rng(31)
    nrow, ncol = 6, 7
    na = O.narcs(nrow, ncol)
    w = rng.uniform(0.05, 0.5, na)
    phi = rng.uniform(0, TWOPI, (nrow, ncol))
    truth = rng.normal(0, 1, (nrow, ncol))
    t, h = arc_list(nrow, ncol)
    g = wrap_pi(phi.ravel()[h] - phi.ravel()[t])

    def dense_u(wv):
        L = dense_laplacian(wv, nrow, ncol)
        b = np.zeros(nrow * ncol)
        np.add.at(b, h, wv * g)
        np.add.at(b, t, -wv * g)
        x = np.linalg.solve(L, b)
        return x - x.mean()

    u, _ = O.wls_solve_f64(w, phi)
    np.testing.assert_allclose(u.ravel(), dense_u(w), atol=1e-9)
    loss = lambda uu: 0.5 * np.sum((uu - (truth.ravel() - truth.mean())) ** 2)  # noqa: E731
    gu = (u.ravel() - (truth.ravel() - truth.mean())).reshape(nrow, ncol)
    gw, _ = O.wls_adjoint_f64(w, phi, u, gu)
    for a in rng.choice(na, 12, replace=False):
        e = np.zeros(na)
        e[a] = 1e-6
        fd = (loss(dense_u(w + e)) - loss(dense_u(w - e))) / 2e-6
        assert gw[a] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_synth_generator_frozen():
    s = synth.make_scene(64, 80, nbumps=3, npatches=1, nlooks=5.0, seed=7, patch=20)
    assert synth.checksum(s.igram, s.corr) == GOLDEN_SYNTH_SHA


GOLDEN_SYNTH_SHA = "062084a27175bc59f7dc1561622450e223c4eca21a96c51ca6d308ee015ac156"


# ── MCF path: incremental costs and the cost-gated conncomp ────────────────
def test_incrcost_smooth_known_answers():
    # rows: (offset, sigsq, flow, nflow) -> (poscost, negcost), nshortcycle 200
    cases = [
        ((0, 32000, 5, 1), (0, 0)),            # LARGESHORT: no preference
        ((0, 100, 0, 1), (400, 400)),          # 200^2/100 both ways
        ((-200, 100, 1, 1), (400, 400)),       # a unit flow that fits exactly
        ((150, 100, 0, 1), (1000, -200)),      # 1225-225, 25-225
        ((150, 100, 0, 2), (700, 100)),        # ceil(2800/4), ceil(400/4)
        ((0, 1, 100, 1), (32000, -32000)),     # clamped both ways
        ((-7, 4, 0, 2), (9650, 10350)),        # (393^2//4 - 12)/4, (407^2//4 - 12)/4
        ((-7, 3, 0, 2), (12867, 13800)),       # ceil(51467/4) = 12867, 55200/4
        ((390, 3, 0, 2), (32000, -12667)),     # ceil(157333/4) clamped; floor(-50667/4)
    ]
    for (off, sig, f, nfl), exp in cases:
        pc, nc, _ = O.incrcost_smooth([[off, sig]], [f], nshortcycle=200, nflow=nfl)
        assert (int(pc[0]), int(nc[0])) == exp, (off, sig, f, nfl)


def test_incrcost_smooth_exact_integer_arithmetic_and_early_exit():
    rng = np.random.default_rng(3)
    n = 4000
    off = rng.integers(-3000, 3000, n)
    sig = rng.integers(1, 32000, n)
    sig[::97] = 32000
    flo = rng.integers(-30, 30, n)
    for nfl in (1, 2, 3):
        pc, nc, ee = O.incrcost_smooth(np.stack([off, sig], 1), flo, nshortcycle=200, nflow=nfl)
        for i in range(0, n, 7):
            if sig[i] == 32000:
                assert pc[i] == 0 and nc[i] == 0
                continue
            c1 = abs(int(flo[i]) * 200 + int(off[i])) ** 2 // int(sig[i])
            for k, got in ((nfl, pc[i]), (-nfl, nc[i])):
                raw = abs((int(flo[i]) + k) * 200 + int(off[i])) ** 2 // int(sig[i]) - c1
                q = -(-raw // (nfl * nfl)) if raw > 0 else raw // (nfl * nfl)
                assert int(got) == max(-32000, min(32000, q))
        assert ee == bool(pc.min() >= 0 and nc.min() >= 0)
    # zero flows on zero offsets: every incremental cost is >= 0 -> early exit
    pc, nc, ee = O.incrcost_smooth(np.stack([np.zeros(50, int), np.full(50, 300)], 1), np.zeros(50))
    assert ee and pc.min() > 0


def test_conncomp_gate_splits_on_low_incremental_cost():
    nrow, ncol = 3, 4
    unw = np.zeros((nrow, ncol), np.float32)
    corr = np.ones((nrow, ncol), np.float32)
    na = O.narcs(nrow, ncol)
    pos = np.full(na, 1000, np.int16)
    neg = np.full(na, 1000, np.int16)
    nra = (nrow - 1) * ncol
    # cut every arc between column 1 and column 2
    for r in range(nrow):
        neg[nra + r * (ncol - 1) + 1] = 10
    lab = O.conncomp(unw, corr, minfrac=0.0, maxncomps=10, poscost=pos, negcost=neg, conncompthresh=300)
    assert (lab[:, :2] == 1).all() and (lab[:, 2:] == 2).all()
    # gate off (costs above threshold): one component; no costs: one component
    assert (O.conncomp(unw, corr, minfrac=0.0, poscost=pos, negcost=pos, conncompthresh=300) == 1).all()
    assert (O.conncomp(unw, corr, minfrac=0.0) == 1).all()


def test_conncomp_gated_matches_scipy_partition():
    rng = np.random.default_rng(12)
    nrow, ncol = 40, 55
    unw = np.cumsum(rng.normal(0, 1.0, (nrow, ncol)), axis=1).astype(np.float32)
    corr = rng.uniform(0, 1, (nrow, ncol)).astype(np.float32)
    na = O.narcs(nrow, ncol)
    pos = rng.integers(0, 800, na).astype(np.int16)
    neg = rng.integers(0, 800, na).astype(np.int16)
    lab = O.conncomp(unw, corr, minfrac=0.0, maxncomps=100000, poscost=pos, negcost=neg, conncompthresh=300)
    valid = ~(corr < 0.05)
    nra = (nrow - 1) * ncol
    ok = np.minimum(pos, neg) >= 300
    v_ok = valid[:-1, :] & valid[1:, :] & (np.abs(unw[:-1, :] - unw[1:, :]) < np.float32(np.pi)) & \
        ok[:nra].reshape(nrow - 1, ncol)
    h_ok = valid[:, :-1] & valid[:, 1:] & (np.abs(unw[:, :-1] - unw[:, 1:]) < np.float32(np.pi)) & \
        ok[nra:].reshape(nrow, ncol - 1)
    idx = np.arange(nrow * ncol).reshape(nrow, ncol)
    ri = np.concatenate([idx[:, :-1][h_ok], idx[:-1, :][v_ok]])
    ci = np.concatenate([idx[:, 1:][h_ok], idx[1:, :][v_ok]])
    g = coo_matrix((np.ones(ri.size), (ri, ci)), shape=(nrow * ncol, nrow * ncol))
    _, comp = connected_components(g, directed=False)
    comp = comp.reshape(nrow, ncol)
    # same partition of the valid pixels (labels are a relabelling)
    pairs = set(zip(comp[valid].tolist(), lab[valid].tolist()))
    assert len(pairs) == len(set(comp[valid].tolist())) == len(set(lab[valid].tolist()))


# ── post-passes: the reference's own hook tests (test_fix_cycle_spikes.py,
#    test_laplace_neighbor_feedback.py) run on the oracle ──────────────────────

def _ramp(nrow, ncol, a, b):
    r, c = np.mgrid[0:nrow, 0:ncol]
    return a * r + b * c


def test_oracle_fix_cycle_spikes_reference_cases():
    two_pi = 2.0 * np.pi
    ramp = _ramp(60, 40, 0.02, 0.01)
    u = ramp.astype(np.float32).copy()
    u[27, :] -= np.float32(two_pi)   # isolated row
    np.testing.assert_allclose(O.fix_cycle_spikes(u), ramp, atol=1e-3)
    ramp = _ramp(40, 60, 0.015, 0.02)
    u = ramp.astype(np.float32).copy()
    u[:, 33] += np.float32(2.0 * two_pi)   # isolated column, two cycles
    np.testing.assert_allclose(O.fix_cycle_spikes(u), ramp, atol=1e-3)
    u = _ramp(50, 50, 0.05, 0.03).astype(np.float32)   # smooth ramp: untouched
    np.testing.assert_array_equal(O.fix_cycle_spikes(u), u)
    ramp = _ramp(60, 40, 0.02, 0.01)
    u = ramp.astype(np.float32).copy()
    u[20:40, :] -= np.float32(two_pi)   # sustained band: untouched
    np.testing.assert_array_equal(O.fix_cycle_spikes(u), u)
    u = ramp.astype(np.float32).copy()
    u[27, :] -= np.float32(two_pi)
    m = np.ones((60, 40), np.uint8)
    m[27, :10] = 0   # masked pixels neither sampled nor corrected
    out = O.fix_cycle_spikes(u, m)
    np.testing.assert_array_equal(out[27, :10], u[27, :10])
    np.testing.assert_allclose(out[27, 10:], ramp[27, 10:], atol=1e-3)


def test_oracle_fix_cycle_spikes_short_lines_and_edges():
    """Lines with <= 20 valid pixels have no median and are never judged;
    the first and last lines are never shifted."""
    u = _ramp(30, 20, 0.02, 0.01).astype(np.float32)   # rows of 20: no row medians
    u[10, :] -= np.float32(2 * np.pi)
    out = O.fix_cycle_spikes(u)
    np.testing.assert_array_equal(out[10, :], u[10, :])   # rows skipped (20 valid)
    u = _ramp(40, 40, 0.02, 0.01).astype(np.float32)
    u[0, :] -= np.float32(2 * np.pi)
    np.testing.assert_array_equal(O.fix_cycle_spikes(u)[0], u[0])


def test_oracle_neighbor_feedback_reference_cases():
    unw = np.zeros((50, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    m = np.ones((50, 200), np.uint8)
    fixed = O.neighbor_feedback(unw, m, 1, 2, 0, 64, 50)
    assert abs(fixed[0, 100] - fixed[0, 99]) < 0.5
    np.testing.assert_allclose(fixed[:, :90], unw[:, :90])
    np.testing.assert_allclose(fixed[:, 160:], unw[:, 160:])
    # closed form: seam at 1*(200-64)/2 + 32 = 100, diff = -5 on every row,
    # unw[:, c] += -5 * (1 - (c - 100)/50) for c in [100, 150)
    c = np.arange(100, 150)
    w = (np.float32(1) - (c - 100).astype(np.float32) / np.float32(50)).astype(np.float32)
    expect = (np.float32(15) + (-5.0 * w.astype(np.float64)).astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(fixed[:, 100:150], np.broadcast_to(expect, (50, 50)))
    # local gradient: the seam mismatch must not grow
    r, cc = np.mgrid[0:500, 0:200]
    unw = (0.15 * cc).astype(np.float32)
    before = np.abs(unw[:, 100] - unw[:, 99]).mean()
    fixed = O.neighbor_feedback(unw, np.ones((500, 200), np.uint8), 1, 2, 0, 64, 200)
    assert np.abs(fixed[:, 100] - fixed[:, 99]).mean() <= before * 1.1
    # sparse boundary data is interpolated, not skipped
    unw = np.zeros((60, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    m = np.ones((60, 200), np.uint8)
    m[:, 99:101] = 0
    m[::10, 99:101] = 1
    fixed = O.neighbor_feedback(unw, m, 1, 2, 0, 64, 50)
    corr = fixed[:, 110] - unw[:, 110]
    sparse = np.delete(corr, np.arange(0, 60, 10))
    assert np.abs(sparse).min() > 1.0
    np.testing.assert_allclose(sparse, np.median(corr[::10]), atol=1.0)
    assert np.abs(np.diff(corr)).max() < 2.0
    # far from any measurement the correction fades to zero
    unw = np.zeros((400, 200), np.float32)
    unw[:, :100], unw[:, 100:] = 10.0, 15.0
    m = np.zeros((400, 200), np.uint8)
    m[:, 150] = 1
    m[190:200, 99:101] = 1
    corr = O.neighbor_feedback(unw, m, 1, 2, 0, 64, 50)[:, 150] - unw[:, 150]
    for i in (0, -1, 50, 350):
        assert abs(corr[i]) < 0.5
    # single tile: no seam
    unw = np.full((40, 40), 3.5, np.float32)
    np.testing.assert_array_equal(O.neighbor_feedback(unw, None, 1, 1, 0, 0, 200), unw)


def test_oracle_neighbor_feedback_fill_closed_forms():
    """Two measured rows 10 apart (gap <= 2*42): the moving median of the two
    is the upper one (rank 1 of 2) at both, and rows between are linear in
    between (equal here); beyond the last measurement the profile fades
    linearly to 0 over 42 rows."""
    nrow, ncol = 200, 64
    unw = np.zeros((nrow, ncol), np.float32)
    unw[:, 32:] = 4.0
    m = np.zeros((nrow, ncol), np.uint8)
    m[:, 40] = 1
    m[[50, 60], 31:33] = 1
    out = O.neighbor_feedback(unw, m, 1, 2, 0, 0, 16)   # seam at 64/2 = 32
    d = (out[:, 40] - unw[:, 40]).astype(np.float64)
    w40 = np.float32(1) - np.float32(8) / np.float32(16)
    np.testing.assert_allclose(d[50:61], -4.0 * w40, rtol=0, atol=1e-6)
    for k in (1, 20, 41):   # fade after row 60 and before row 50
        np.testing.assert_allclose(d[60 + k], -4.0 * (1 - k / 42) * w40, atol=1e-5)
        np.testing.assert_allclose(d[50 - k], -4.0 * (1 - k / 42) * w40, atol=1e-5)
    assert d[60 + 42] == 0.0 and d[50 - 42] == 0.0


# ── phase bridging: the reference's hook tests (test_bridge.py) and its stages
#    against scipy.ndimage ────────────────────────────────────────────────────

def _bridge_fixture_three_regions():
    nrow, ncol = 60, 90
    rng = np.random.default_rng(0)
    r, c = np.mgrid[0:nrow, 0:ncol]
    ramp = 0.02 * r + 0.015 * c
    off = np.zeros((nrow, ncol))
    off[(r >= 20) & (r < 39)] = 3.0 * 2 * np.pi
    off[r >= 40] = -2.0 * 2 * np.pi
    unw = (ramp + off).astype(np.float32) + rng.normal(0.0, 0.05, (nrow, ncol)).astype(np.float32)
    m = np.ones((nrow, ncol), np.uint8)
    m[19, :] = 0
    m[39, :] = 0
    return unw, m, ramp


def test_oracle_bridge_reference_cases():
    two_pi = 2 * np.pi
    unw, m, ramp = _bridge_fixture_three_regions()
    u, cc, n = O.bridge(unw, m, min_num_pixel=1, erosion_size=0)
    v = m == 1
    assert n == 3 and (u[v] - ramp[v]).std() < 0.3 and not np.allclose(u[v], unw[v])
    u, cc, n = O.bridge(np.full((30, 30), 1.2345, np.float32))
    assert n == 1 and np.array_equal(u, np.full((30, 30), 1.2345, np.float32)) and not cc.any()
    r, c = np.mgrid[0:40, 0:60]
    ramp = 1.0 + 0.02 * r + 0.01 * c
    unw = (ramp + np.where(c >= 30, 4.0 * two_pi, 0.0)).astype(np.float32)
    m = np.ones((40, 60), np.uint8)
    m[:, 28:32] = 0
    unw[:, 28:32] = 99.0
    u, cc, n = O.bridge(unw, m, min_num_pixel=1, erosion_size=0)
    v = m == 1
    assert (u[v] - ramp[v]).std() < 0.3
    np.testing.assert_array_equal(u[:, 28:32], unw[:, 28:32])
    r, _ = np.mgrid[0:200, 0:10]
    ramp = 5.0 + 0.2 * r.astype(np.float64)
    unw = ramp.astype(np.float32)
    m = np.ones((200, 10), np.uint8)
    m[90:110, :] = 0
    u0, _, _ = O.bridge(unw, m, radius=15, min_num_pixel=1, erosion_size=0)
    u1, _, _ = O.bridge(unw, m, radius=15, min_num_pixel=1, erosion_size=0, ramp_type="linear_azimuth")
    v = m == 1
    assert (u0[v] - ramp[v]).std() > 1.0
    np.testing.assert_allclose(u1[v], ramp[v], atol=1e-3)


def test_oracle_bridge_stages_match_scipy():
    from scipy import ndimage

    rng = np.random.default_rng(11)
    topo = ndimage.gaussian_filter(rng.normal(size=(120, 140)), 3)
    topo = np.where(topo > 0.05, 1.0 + rng.random((120, 140)), 0.0).astype(np.float32)
    lab, n = O.bridge_label(topo)
    ref, nref = ndimage.label(topo != 0, structure=np.ones((3, 3)))
    assert n == nref > 3
    np.testing.assert_array_equal(lab, ref.astype(np.uint32))
    for e in (1, 2, 3):
        for circ in (0, 1):
            se = np.ones((2 * e + 1, 2 * e + 1), bool)
            if circ:
                dr, dc = np.mgrid[-e:e + 1, -e:e + 1]
                se = dr * dr + dc * dc <= e * e
            er = ndimage.binary_erosion(lab > 0, structure=se, border_value=0)
            keep = np.unique(lab[er])
            keep = keep[keep > 0]
            remap = np.zeros(n + 1, np.uint32)
            remap[keep] = np.arange(1, len(keep) + 1)
            got, k = O.bridge_erode_prune(lab, n, e, circ)
            assert k == len(keep)
            np.testing.assert_array_equal(got, remap[lab])
    b = O.bridge_boundary(lab)
    mx = ndimage.maximum_filter(lab, size=3, mode="reflect")
    mn = ndimage.minimum_filter(lab, size=3, mode="reflect")
    np.testing.assert_array_equal(b.astype(bool), (mx != mn) & (lab > 0))


def test_oracle_bridge_nearest_pairs_brute_force():
    """Per ordered region pair: the minimum of (floor(64*dist), local index
    of the first point, of the second) over the subsampled boundary points,
    by a numpy brute force."""
    rng = np.random.default_rng(12)
    lab = np.zeros((50, 70), np.uint32)
    lab[2:15, 3:30], lab[20:45, 5:20], lab[25:48, 40:66], lab[3:12, 45:60] = 1, 2, 3, 4
    b = O.bridge_boundary(lab)
    for ms in (4096, 7):
        dm, ep = O.bridge_nn(lab, b, 4, ms)
        pts = {}
        for l in range(1, 5):
            idx = np.flatnonzero((b.ravel() > 0) & (lab.ravel() == l))
            stride = max(1, len(idx) // ms)
            pts[l] = idx[::stride][:ms]
        for i in range(1, 5):
            for j in range(1, 5):
                if i == j:
                    assert dm[i, j] == -1
                    continue
                pi, pj = pts[i], pts[j]
                dy = (pj[None, :] // 70 - pi[:, None] // 70).astype(np.float32)
                dx = (pj[None, :] % 70 - pi[:, None] % 70).astype(np.float32)
                d = np.sqrt(dx * dx + dy * dy).astype(np.float32)
                key = (np.floor(d * np.float32(64)).astype(np.uint64) << np.uint64(24)) | \
                      (np.arange(len(pi), dtype=np.uint64)[:, None] << np.uint64(12)) | \
                      np.arange(len(pj), dtype=np.uint64)[None, :]
                a, bb = np.unravel_index(np.argmin(key), key.shape)
                assert dm[i, j] == np.float32(int(key[a, bb] >> np.uint64(24))) / np.float32(64)
                assert tuple(ep[i, j]) == (pi[a] // 70, pi[a] % 70, pj[bb] // 70, pj[bb] % 70)

"""MCF path's GPU stages (incremental costs, cost-gated connected components,
flow integration) and the host network-flow solver hook, against the oracle.
The network-flow solve itself (SNAPHU's CPU code in the reference) is not
part of this library; the tests register stand-in solvers whose flows are
known: zero flows on a residue-free ramp (the optimum there) and a fixed
integer flow field."""

import ctypes as C

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi, synth
from cuphu_amd.mcf import incremental_costs
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


@pytest.fixture
def solver_slot():
    yield
    cuphu.set_flow_solver(None)


@pytest.mark.parametrize("nflow", [1, 2])
def test_incremental_costs_bitexact(gpu_ready, nflow):
    torch = G.torch_mod()
    rng = np.random.default_rng(nflow)
    n = 200_003
    costs = np.stack([rng.integers(-3000, 3000, n), rng.integers(1, 32000, n)], 1).astype(np.int16)
    costs[::101, 1] = 32000
    flows = rng.integers(-20, 20, n).astype(np.int16)
    dc, df = G.dev(costs), G.dev(flows)
    pos, neg, flag = incremental_costs(dc, df, nshortcycle=200, nflow=nflow)
    torch.cuda.synchronize()
    rp, rn, rflag = O.incrcost_smooth(costs, flows, nshortcycle=200, nflow=nflow)
    np.testing.assert_array_equal(G.host(pos), rp)
    np.testing.assert_array_equal(G.host(neg), rn)
    assert flag == rflag
    # the early-exit flag on an all-non-negative case
    z = np.stack([np.zeros(1000, np.int16), np.full(1000, 300, np.int16)], 1)
    _, _, f2 = incremental_costs(G.dev(z), G.dev(np.zeros(1000, np.int16)))
    assert f2


@pytest.mark.parametrize("nshortcycle,nflow", [(200, 1), (200, 3), (5000, 2)])
@pytest.mark.parametrize("aligned", [True, False])
def test_incremental_costs_full_range_bitexact(gpu_ready, nshortcycle, nflow, aligned):
    """Offsets, flows and sigsq over the whole int16 range (sigsq >= 1, plus
    the masked sentinel), a fifth of the flows small.  At nshortcycle 5000
    most arcs have |flow * nshortcycle + offset| >= 2^25 and take the 64-bit
    integer division instead of the exact FP64 form.  Unaligned: the
    inputs start one element into their buffers (the scalar-access kernel)."""
    torch = G.torch_mod()
    rng = np.random.default_rng(nshortcycle + nflow)
    n = 300_007
    costs = np.stack([rng.integers(-32768, 32768, n), rng.integers(1, 32768, n)], 1).astype(np.int16)
    costs[::97, 1] = 32000
    flows = rng.integers(-32768, 32768, n).astype(np.int16)
    flows[::5] = rng.integers(-3, 4, flows[::5].size)
    k = 0 if aligned else 1
    dc = G.dev(np.concatenate([np.zeros((k, 2), np.int16), costs]))[k:]
    df = G.dev(np.concatenate([np.zeros(k, np.int16), flows]))[k:]
    assert (dc.data_ptr() % 16 == 0) == aligned and (df.data_ptr() % 8 == 0) == aligned
    pos, neg, flag = incremental_costs(dc, df, nshortcycle=nshortcycle, nflow=nflow)
    torch.cuda.synchronize()
    rp, rn, rflag = O.incrcost_smooth(costs, flows, nshortcycle=nshortcycle, nflow=nflow)
    np.testing.assert_array_equal(G.host(pos), rp)
    np.testing.assert_array_equal(G.host(neg), rn)
    assert flag == rflag


@pytest.mark.parametrize("shape", [(23, 31), (130, 300)])
def test_conncomp_gated_bitexact(gpu_ready, shape):
    torch = G.torch_mod()
    rng = np.random.default_rng(shape[0])
    unw = np.cumsum(rng.normal(0, 0.9, shape), axis=1).astype(np.float32)
    corr = rng.uniform(0, 1, shape).astype(np.float32)
    na = O.narcs(*shape)
    pos = rng.integers(0, 900, na).astype(np.int16)
    neg = rng.integers(0, 900, na).astype(np.int16)
    prm = _capi.default_params()
    prm.minconncompfrac = 0.0
    prm.maxncomps = 100000
    du, dc, dp, dn = G.dev(unw), G.dev(corr), G.dev(pos), G.dev(neg)
    lab = G.empty(shape, torch.int32)
    _capi.check(_capi.lib.cuphu_conncomp_gpu(G.p(du), G.p(dc), None, shape[0], shape[1], G.p(dp), G.p(dn),
                                             C.byref(prm), 0, G.p(lab)), "conncomp")
    ref = O.conncomp(unw, corr, None, 0.0, 100000, poscost=pos, negcost=neg, conncompthresh=prm.conncompthresh)
    np.testing.assert_array_equal(G.host(lab).view(np.uint32), ref)


def test_mcf_zero_flows_on_ramp_is_closed_form(gpu_ready, solver_slot):
    """Config 1's 'mcf' leg: the 256x256 ramp has no residues, so zero flows
    are the network-flow optimum; the GPU stages must then give the closed
    form unw = pi(x+y) + 6pi with one component."""
    s = synth.ramp_256()
    seen = {}

    def zero(costs, phase, init):
        seen.update(costs=costs.shape, phase=phase.shape, init=init)
        return np.zeros(costs.shape[0], np.int16)

    cuphu.set_flow_solver(zero)
    unw, cc = cuphu.unwrap(s.igram, s.corr, 1.0, init="mcf")
    assert seen == {"costs": (O.narcs(256, 256), 2), "phase": (256, 256), "init": "mcf"}
    assert np.abs(unw - s.truth).max() < 1e-3
    assert (cc == 1).all()


def test_mcf_flows_integrated_and_conncomp_gated(gpu_ready, solver_slot):
    """A fixed integer flow field: the unwrapped phase is the oracle's
    integration of those flows, and the components are the oracle's with
    the incremental-cost gate at those flows."""
    s = synth.make_scene(90, 120, nbumps=2, npatches=1, nlooks=5.0, seed=8, patch=25)
    nrow, ncol = s.corr.shape
    na = O.narcs(nrow, ncol)
    rng = np.random.default_rng(5)
    flows = np.where(rng.uniform(size=na) < 0.02, rng.integers(-2, 3, na), 0).astype(np.int16)
    got = {}

    def fixed(costs, phase, init):
        got["costs"] = costs
        got["phase"] = phase
        return flows

    cuphu.set_flow_solver(fixed)
    unw, cc = cuphu.unwrap(s.igram, s.corr, 5.0, cost="smooth", init="mcf", reference_pixel=None,
                           min_conncomp_frac=0.0)
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    np.testing.assert_array_equal(got["phase"], ph)
    d = np.abs(got["costs"].astype(np.int32) - O.costs(ph, s.corr, 5.0))
    assert d.max() <= 1 and (d > 0).sum() <= 2
    nra = (nrow - 1) * ncol
    ref_unw = O.integrate_phase(ph, flows[:nra].reshape(nrow - 1, ncol), flows[nra:].reshape(nrow, ncol - 1))
    np.testing.assert_array_equal(unw, ref_unw)
    prm = _capi.default_params()
    pc, nc, _ = O.incrcost_smooth(got["costs"], flows, nshortcycle=int(prm.nshortcycle))
    ref_cc = O.conncomp(ref_unw, s.corr, None, 0.0, prm.maxncomps, poscost=pc, negcost=nc,
                        conncompthresh=prm.conncompthresh)
    np.testing.assert_array_equal(cc, ref_cc)
    assert cc.max() >= 1


def test_mcf_defo_cost_and_mst_reach_the_solver(gpu_ready, solver_slot):
    s = synth.make_scene(40, 50, nbumps=1, npatches=0, nlooks=10.0, seed=9)
    seen = []
    cuphu.set_flow_solver(lambda c, p, i: (seen.append((c.shape[1], i)), np.zeros(c.shape[0], np.int16))[1])
    cuphu.unwrap(s.igram, s.corr, 10.0, cost="defo", init="mst")
    assert seen == [(4, "mst")]


def test_mcf_without_solver_raises(gpu_ready, solver_slot):
    s = synth.make_scene(40, 50, nbumps=1, npatches=0, nlooks=10.0, seed=9)
    cuphu.set_flow_solver(None)
    with pytest.raises(ValueError, match="network-flow solver"):
        cuphu.unwrap(s.igram, s.corr, 10.0, init="mcf")
    with pytest.raises(ValueError, match="network-flow solver"):
        cuphu.unwrap(s.igram, s.corr, 10.0, init="mcf", ntiles=(2, 2))


def _tile_flows(phase, seed):
    """A deterministic integer flow field for one tile (about 2 % of arcs
    nonzero, in -2..2), from the tile's own shape and a seed."""
    nrow, ncol = phase.shape
    na = O.narcs(nrow, ncol)
    rng = np.random.default_rng(seed * 100003 + nrow * 1009 + ncol)
    return np.where(rng.uniform(size=na) < 0.02, rng.integers(-2, 3, na), 0).astype(np.int16)


@pytest.mark.parametrize("ntiles,overlap,masked", [((2, 2), 24, False), ((2, 3), (20, 30), True), ((3, 1), 0, False)])
def test_tiled_mcf_matches_oracle_per_tile_and_integer_stitch(gpu_ready, solver_slot, ntiles, overlap, masked):
    """init='mcf' on a tiled scene (cuphu_solver.cpp:1753-1808, :1928): each
    tile of the reference tile formula goes to the registered solver with
    its own phase and costs; the result is the oracle's integration of each
    tile's flows, components gated by that tile's incremental costs, tiles
    stitched by whole cycles with the labels concatenated."""
    s = synth.make_scene(150, 190, nbumps=2, npatches=1, nlooks=5.0, seed=12, patch=40)
    nrow, ncol = s.corr.shape
    mask = None
    if masked:
        mask = np.ones((nrow, ncol), np.uint8)
        mask[60:75, 20:90] = 0
    calls = []

    def solver(costs, phase, init):
        f = _tile_flows(phase, len(calls))
        calls.append((costs.copy(), phase.copy(), f))
        return f

    cuphu.set_flow_solver(solver)
    unw, cc = cuphu.unwrap(s.igram, s.corr, 5.0, cost="smooth", init="mcf", ntiles=ntiles, tile_overlap=overlap,
                           mask=mask, reference_pixel=None, min_conncomp_frac=0.0)
    rov, cov = (overlap, overlap) if np.ndim(overlap) == 0 else overlap
    rf, rs = O.tile_axis(nrow, 0, rov, ntiles[0])
    cf, cs = O.tile_axis(ncol, 0, cov, ntiles[1])
    assert len(calls) == len(rf) * len(cf)
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    prm = _capi.default_params()
    tu, tc, tr, tm = [], [], [], []
    for t, (costs, phase, flows) in enumerate(calls):
        r0, c0 = rf[t // len(cf)], cf[t % len(cf)]
        nr, nc = rs[t // len(cf)], cs[t % len(cf)]
        np.testing.assert_array_equal(phase, ph[r0:r0 + nr, c0:c0 + nc])
        crop_corr = s.corr[r0:r0 + nr, c0:c0 + nc]
        crop_mask = None if mask is None else mask[r0:r0 + nr, c0:c0 + nc]
        ref_costs = G.sigsq_from_mask(phase, crop_corr, 5.0, crop_mask)
        d = np.abs(costs.astype(np.int32) - ref_costs)
        assert d.max() <= 1 and (d > 0).sum() <= 2
        nra = (nr - 1) * nc
        u = O.integrate_phase(phase, flows[:nra].reshape(nr - 1, nc), flows[nra:].reshape(nr, nc - 1))
        pc, ng, _ = O.incrcost_smooth(costs, flows, nshortcycle=int(prm.nshortcycle))
        tu.append(u)
        tc.append(O.conncomp(u, crop_corr, crop_mask, 0.0, prm.maxncomps, poscost=pc, negcost=ng,
                             conncompthresh=prm.conncompthresh))
        tr.append(crop_corr)
        tm.append(crop_mask)
    ref_unw, ref_cc, tk = O.stitch(nrow, ncol, rf, rs, cf, cs, rov, cov, tu, tc, tr,
                                   tiles_conf=tm if mask is not None else None, laplace=False)
    assert np.all(tk == np.rint(tk))
    np.testing.assert_array_equal(unw, ref_unw)
    np.testing.assert_array_equal(cc, ref_cc)


def test_mcf_failing_solver_reports_error(gpu_ready, solver_slot):
    s = synth.make_scene(30, 30, nbumps=1, npatches=0, nlooks=10.0, seed=9)

    def bad(costs, phase, init):
        raise RuntimeError("solver failure on purpose")

    cuphu.set_flow_solver(bad)
    with pytest.raises(RuntimeError, match="network-flow solver failed"):
        cuphu.unwrap(s.igram, s.corr, 10.0, init="mcf")


def test_mcf_auto_tiling_with_nproc(gpu_ready, solver_slot):
    """unwrap(init='mcf', nproc > 1) auto-tiles toward target_tile_size (the
    reference front door's rule, _unwrap.py:484-512) and runs each tile
    through the registered solver: the same result as the explicit tile
    counts of that rule."""
    from cuphu_amd._unwrap import _auto_ntiles_by_target_size

    s = synth.make_scene(300, 420, nbumps=2, npatches=0, nlooks=10.0, seed=14)
    counts = []
    cuphu.set_flow_solver(lambda c, p, i: (counts.append(p.shape), np.zeros(c.shape[0], np.int16))[1])
    u1, c1 = cuphu.unwrap(s.igram, s.corr, 10.0, init="mcf", nproc=4, target_tile_size=200)
    ntr, ntc = _auto_ntiles_by_target_size(300, 420, 200, 0, 0)
    assert ntr * ntc > 1 and len(counts) == ntr * ntc
    u2, c2 = cuphu.unwrap(s.igram, s.corr, 10.0, init="mcf", ntiles=(ntr, ntc))
    np.testing.assert_array_equal(u1, u2)
    np.testing.assert_array_equal(c1, c2)

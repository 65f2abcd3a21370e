"""The golden fixture tests/golden/scene64x80.npz (written by
tests/golden/make_fixtures.py from the CPU oracle): the oracle must keep
reproducing it (CPU), and the HIP path is checked against it (GPU) without
running the oracle: arc costs (rounding-boundary +-1 cases counted),
residue charges, the FP64 solve's 2 pi ambiguity counts and the connected
components bit for bit, the FP32 solve's ambiguities on >= 99.9 % of
pixels."""

import os

import numpy as np
import pytest

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene64x80.npz")
TWO_PI = 2.0 * np.pi


@pytest.fixture(scope="module")
def golden():
    with np.load(FIXTURE, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _make():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "make_fixtures", os.path.join(os.path.dirname(FIXTURE), "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_oracle_reproduces_golden_fixture(golden):
    fresh = _make().make()
    for k in ("igram", "corr", "phi", "costs_smooth", "costs_defo", "residues", "amb", "labels"):
        np.testing.assert_array_equal(fresh[k], golden[k], err_msg=k)
    np.testing.assert_allclose(fresh["unw"], golden["unw"], rtol=0, atol=1e-9)


def test_golden_fixture_is_consistent(golden):
    """The stored outputs agree with each other: congruent unwrapped phase,
    ambiguities that match it, label 0 exactly on the sub-floor patches."""
    phi = golden["phi"].astype(np.float64)
    k = (golden["unw"] - phi) / TWO_PI
    np.testing.assert_allclose(k, np.rint(k), atol=1e-6)
    np.testing.assert_array_equal(np.rint(k).astype(np.int32), golden["amb"])
    np.testing.assert_array_equal(golden["labels"] == 0, golden["corr"] < 0.05)


@pytest.mark.gpu
def test_gpu_matches_golden_fixture(gpu_ready, golden):
    import cuphu_amd as cuphu
    from cuphu_amd import _capi
    from cuphu_amd._ext import _cuphu_ext
    from tests import gpu_util as G

    torch = G.torch_mod()
    ig, corr, nlooks = golden["igram"], golden["corr"], float(golden["nlooks"])
    nrow, ncol = corr.shape
    for cost, ne in (("smooth", 2), ("defo", 4)):
        got = _cuphu_ext.build_costs(ig, corr, nlooks, cost=cost).view(np.int16).reshape(-1, ne)
        d = np.abs(got.astype(np.int32) - golden["costs_" + cost])
        print(f"{cost} cost +-1 cases: {int((d > 0).sum())} of {d.size}")
        assert d.max() <= 1 and (d > 0).sum() <= 2
    dphi = G.dev(golden["phi"])
    dres = G.empty((nrow - 1, ncol - 1), torch.int8)
    _capi.check(_capi.lib.cuphu_residues_gpu(G.p(dphi), nrow, ncol, G.p(dres), G.stream()), "residues")
    np.testing.assert_array_equal(G.host(dres), golden["residues"])
    opts = dict(n_irls=int(golden["n_irls"]), max_iter=int(golden["max_iter"]), tol=float(golden["tol"]))
    phi = golden["phi"].astype(np.float64)
    u64, cc64 = cuphu.unwrap(ig, corr, nlooks, cost="defo", init="laplace", ntiles=(1, 1), reference_pixel=None,
                             laplace_options={**opts, "precision": 64})
    np.testing.assert_array_equal(np.rint((u64 - phi) / TWO_PI).astype(np.int32), golden["amb"])
    np.testing.assert_array_equal(cc64, golden["labels"])
    print("FP64 mode: max |unw - golden|", float(np.abs(u64 - golden["unw"]).max()))
    u32, _ = cuphu.unwrap(ig, corr, nlooks, cost="defo", init="laplace", ntiles=(1, 1), reference_pixel=None,
                          laplace_options=opts)
    agree = float((np.rint((u32 - phi) / TWO_PI).astype(np.int32) == golden["amb"]).mean())
    print(f"FP32: ambiguity agreement {agree:.6f}, max |unw - golden| {float(np.abs(u32 - golden['unw']).max()):.3e}")
    assert agree >= 0.999

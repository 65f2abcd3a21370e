"""Parity at the workloads' own sizes (GPU) against the CPU FP64 oracle.

* 2048x2048 interferogram (ramp + 5 bumps + one 300x300 decorrelated patch,
  nlooks 10, defo cost, tol 1e-6, 5 IRLS re-solves, the 1000-iteration
  cap): the fused and plain solvers and the public unwrap().
* The first four interferograms of the 64-instance uneven batch at full
  size (2320x1727 ... 2023x876, per-instance nlooks), solved together in
  one launch sequence.
* Two tiles of the 8192x8192 auto-tiled scene (the 1024x1024 corner tile
  and the 2048x2048 tile at the origin): the FP64 mode, which the benchmark
  uses as its accuracy reference at full size, and the FP32 fused solver.
* A converged solve (tol 1e-10, cap 20000) of a 512x512 crop in the FP64
  mode, the reference for the benchmark's capped-versus-converged figure.

The oracle runs the same algorithm and schedule in FP64 (OpenMP).  The
measure is the 2-pi ambiguity count k = rint((u - phi) / 2 pi): FP32
results must agree on >= 99.9 % of pixels, FP64 modes on >= 99.99 %.
Each test prints the agreement and the max / rms of the unwrapped-phase
difference.
"""

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import _capi, synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi
SCHEDULE = dict(n_irls=5, max_iter=1000, tol=1e-6)


def _report(tag, u, ref, ph):
    ph = ph.astype(np.float64)
    ka = np.rint((u.astype(np.float64) - ph) / TWOPI)
    kb = np.rint((ref.astype(np.float64) - ph) / TWOPI)
    agree = float((ka == kb).mean())
    d = u.astype(np.float64) - ref
    print(f"[{tag}] ambiguity agreement {agree:.7f}, max |du| {np.abs(d).max():.3e} rad, "
          f"rms {np.sqrt(np.mean(d * d)):.3e} rad", flush=True)
    return agree


def _oracle(s, **opts):
    ph = O.wrapped_phase(s.igram, correctly_rounded=True)
    sig = O.costs(ph, s.corr, s.nlooks, defo=True)[:, 1]
    u, st = O.pcg_irls_f64(sig, ph, **opts)
    return u, st, ph


@pytest.fixture(scope="module")
def scene2048():
    s = synth.config_scene("scene2048")
    u, st, ph = _oracle(s, **SCHEDULE)
    print(f"[scene2048] oracle iterations {st['iters']} relres {[f'{x:.2e}' for x in st['relres']]}", flush=True)
    return s, u, ph


@pytest.mark.parametrize("solver", ["fused", "plain"])
def test_scene2048_matches_fp64_oracle(gpu_ready, scene2048, solver):
    s, ref, ph = scene2048
    u, cc, st = G.batch_solve([s], _capi.default_options(solver=solver, **SCHEDULE), cost=_capi.COST["defo"])
    assert st[0]["npass"] == 6
    assert _report(f"scene2048 {solver}", u[0], ref, ph) >= 0.999


def test_scene2048_public_unwrap_matches_fp64_oracle(gpu_ready, scene2048):
    s, ref, ph = scene2048
    unw, cc = cuphu.unwrap(s.igram, s.corr, s.nlooks, cost="defo", init="laplace", ntiles=(1, 1),
                           laplace_options=dict(SCHEDULE))
    assert _report("scene2048 unwrap()", unw, ref, ph) >= 0.999
    prm = _capi.default_params()
    np.testing.assert_array_equal(cc, O.conncomp(unw, s.corr, None, prm.minconncompfrac, prm.maxncomps))


@pytest.fixture(scope="module")
def batch_first4():
    scenes = synth.batch_instances(64, first=4)
    u, _, st = G.batch_solve(scenes, _capi.default_options(**SCHEDULE), want_cc=False, cost=_capi.COST["defo"])
    return scenes, u, st


@pytest.mark.parametrize("i", range(4))
def test_batch64_instances_match_fp64_oracle(gpu_ready, batch_first4, i):
    scenes, u, st = batch_first4
    s = scenes[i]
    ref, sto, ph = _oracle(s, **SCHEDULE)
    print(f"[batch64 #{i}] {s.corr.shape} nlooks {s.nlooks:g}: gpu iters {st[i]['iters']}, oracle {sto['iters']}")
    assert _report(f"batch64 #{i}", u[i], ref, ph) >= 0.999


@pytest.fixture(scope="module")
def scene8192_tiles():
    s = synth.config_scene("scene8192")
    tiles = {}
    for name, (r0, c0, n) in {"corner1024": (7168, 7168, 1024), "origin2048": (0, 0, 2048)}.items():
        tiles[name] = synth.Scene(igram=np.ascontiguousarray(s.igram[r0:r0 + n, c0:c0 + n]),
                                  corr=np.ascontiguousarray(s.corr[r0:r0 + n, c0:c0 + n]),
                                  truth=s.truth[r0:r0 + n, c0:c0 + n], nlooks=s.nlooks, seed=s.seed)
    del s
    return tiles


@pytest.mark.parametrize("tile", ["corner1024", "origin2048"])
def test_scene8192_tile_fp64_mode_and_fp32_match_fp64_oracle(gpu_ready, scene8192_tiles, tile):
    s = scene8192_tiles[tile]
    ref, sto, ph = _oracle(s, **SCHEDULE)
    print(f"[scene8192 {tile}] oracle iterations {sto['iters']}")
    u64, _, st64 = G.batch_solve([s], _capi.default_options(precision=64, **SCHEDULE), want_cc=False,
                                 cost=_capi.COST["defo"])
    assert st64[0]["iters"] == sto["iters"]
    assert _report(f"scene8192 {tile} fp64 mode", u64[0], ref, ph) >= 0.9999
    u32, _, _ = G.batch_solve([s], _capi.default_options(**SCHEDULE), want_cc=False, cost=_capi.COST["defo"])
    assert _report(f"scene8192 {tile} fp32 fused", u32[0], ref, ph) >= 0.999


def test_converged_fp64_mode_matches_converged_oracle(gpu_ready):
    """tol 1e-10 with a 20000-iteration cap: both solves converge on every
    pass, and the FP64 mode's result is the oracle's."""
    s0 = synth.config_scene("scene2048")
    n = 512
    s = synth.Scene(igram=np.ascontiguousarray(s0.igram[:n, :n]), corr=np.ascontiguousarray(s0.corr[:n, :n]),
                    truth=s0.truth[:n, :n], nlooks=s0.nlooks, seed=s0.seed)
    conv = dict(n_irls=5, max_iter=20000, tol=1e-10)
    ref, sto, ph = _oracle(s, **conv)
    assert all(x == 0 for x in sto["status"]), sto
    u64, _, st64 = G.batch_solve([s], _capi.default_options(precision=64, **conv), want_cc=False,
                                 cost=_capi.COST["defo"])
    print(f"[converged 512] gpu iters {st64[0]['iters']}, oracle {sto['iters']}")
    assert all(x == 0 for x in st64[0]["pass_status"]), st64[0]
    assert _report("converged 512 fp64 mode", u64[0], ref, ph) >= 0.9999

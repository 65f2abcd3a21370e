"""unwrap_batch: many uneven interferograms with their own nlooks in one
launch sequence -- bit-identical to solving each alone, agreeing with the
FP64 oracle per instance, with input validation."""

import numpy as np
import pytest

import cuphu_amd as cuphu
import oracle as O
from cuphu_amd import synth

pytestmark = pytest.mark.gpu
TWOPI = 2 * np.pi


def _scenes(n, seed=40):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nr, nc = int(rng.integers(40, 160)), int(rng.integers(40, 200))
        nl = float(rng.choice([1, 5, 10, 20]))
        out.append(synth.make_scene(nr, nc, nbumps=2, npatches=int(i % 2), nlooks=nl, seed=seed + i,
                                    patch=min(15, nr // 3, nc // 3)))
    return out


def test_batch_equals_solo_and_oracle(gpu_ready):
    scenes = _scenes(7)
    opts = dict(n_irls=3, max_iter=500, tol=1e-5)
    unw, cc, st = cuphu.unwrap_batch([s.igram for s in scenes], [s.corr for s in scenes],
                                     [s.nlooks for s in scenes], "defo", laplace_options=opts, return_stats=True)
    assert len(unw) == len(cc) == len(st) == 7
    for i, s in enumerate(scenes):
        u1, c1 = cuphu.unwrap_batch([s.igram], [s.corr], s.nlooks, "defo", laplace_options=opts)
        np.testing.assert_array_equal(unw[i], u1[0])
        np.testing.assert_array_equal(cc[i], c1[0])
        ph = O.wrapped_phase(s.igram, correctly_rounded=True)
        sig = O.costs(ph, s.corr, s.nlooks, defo=True)[:, 1]
        uo, _ = O.pcg_irls_f64(sig, ph, **opts)
        agree = (np.rint((unw[i] - ph) / TWOPI) == np.rint((uo - ph) / TWOPI)).mean()
        assert agree >= 0.999, (i, agree)
        assert unw[i].dtype == np.float32 and cc[i].dtype == np.uint32


def test_batch_torch_inputs_stay_on_device(gpu_ready):
    import torch

    scenes = _scenes(3, seed=90)
    ig = [torch.from_numpy(s.igram).cuda() for s in scenes]
    co = [torch.from_numpy(s.corr).cuda() for s in scenes]
    unw, cc = cuphu.unwrap_batch(ig, co, 10.0, laplace_options=dict(n_irls=1, max_iter=200))
    assert all(u.is_cuda and u.dtype == torch.float32 for u in unw)
    un, cn = cuphu.unwrap_batch([s.igram for s in scenes], [s.corr for s in scenes], 10.0,
                                laplace_options=dict(n_irls=1, max_iter=200))
    for a, b in zip(unw, un):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_batch_validation(gpu_ready):
    s = _scenes(1)[0]
    with pytest.raises(ValueError):
        cuphu.unwrap_batch([s.igram], [s.corr[:-1]], 10.0)
    with pytest.raises(ValueError):
        cuphu.unwrap_batch([s.igram, s.igram], [s.corr], 10.0)
    with pytest.raises(ValueError):
        cuphu.unwrap_batch([s.igram], [s.corr], [1.0, 2.0])
    assert cuphu.unwrap_batch([], [], 1.0) == ([], [])


@pytest.mark.parametrize("dims", [[(300, 333)], [(517, 1090)], [(40, 64), (200, 700), (129, 1030), (70, 1)]])
def test_tile_widths_bit_identical(gpu_ready, monkeypatch, dims):
    """256- and 128-column tiles (forced with CUPHU_LAPLACE_LPR) and row
    strides rounded to 64 or 256 elements (CUPHU_LAPLACE_STRIDE) give the
    same unwrapped phase, components and iteration counts: the per-pixel
    arithmetic depends on neither, the extra stride columns stay zero, and
    the FP64 partial sums of the groupings round to the same FP32 step
    lengths."""
    scenes = [synth.make_scene(nr, nc, nbumps=2, npatches=1, nlooks=10.0, seed=70 + i,
                               patch=max(1, min(20, nr // 3, nc // 3)))
              for i, (nr, nc) in enumerate(dims)]
    opts = dict(n_irls=2, max_iter=300, tol=1e-6)
    res = {}
    # (lanes per tile row, row stride rounding): 256-column tiles on 64- and
    # 256-element strides, 128-column tiles on 256-element strides
    for lpr, stride in (("64", "64"), ("64", "256"), ("32", "256")):
        monkeypatch.setenv("CUPHU_LAPLACE_LPR", lpr)
        monkeypatch.setenv("CUPHU_LAPLACE_STRIDE", stride)
        res[lpr, stride] = cuphu.unwrap_batch([s.igram for s in scenes], [s.corr for s in scenes], 10.0, "defo",
                                              laplace_options=opts, return_stats=True)
    ref = res["64", "64"]
    for key in (("64", "256"), ("32", "256")):
        for i in range(len(scenes)):
            np.testing.assert_array_equal(ref[0][i], res[key][0][i], err_msg=str(key))
            np.testing.assert_array_equal(ref[1][i], res[key][1][i], err_msg=str(key))
            a, b = ref[2][i], res[key][2][i]
            assert (a["status"], a["iters"], a["pass_status"]) == (b["status"], b["iters"], b["pass_status"])
            np.testing.assert_allclose(a["relres"], b["relres"], rtol=1e-9)


@pytest.mark.parametrize("precision", [32, 64])
def test_merged_tiling_bit_identical(gpu_ready, monkeypatch, precision):
    """Instances 256 k + 64 (128) columns wide run their last 64 (128)
    columns in tall tiles (one workgroup for four (two) 256-column tiles;
    CUPHU_LAPLACE_MERGE=0 turns it off, =2 takes it on this small grid,
    which the default leaves plain): the same unwrapped phase,
    components and iteration counts as the plain 256-column tiling, beside
    instances it leaves alone.  The tall tiles group their rows' FP64
    partial sums differently; the sums round to the same FP32 step
    lengths."""
    dims = [(300, 1843), (130, 4160), (77, 50), (200, 700), (40, 257), (150, 1100), (90, 100)]
    scenes = [synth.make_scene(nr, nc, nbumps=2, npatches=1, nlooks=10.0, seed=90 + i,
                               patch=max(1, min(20, nr // 3, nc // 3)))
              for i, (nr, nc) in enumerate(dims)]
    opts = dict(n_irls=2, max_iter=300, tol=1e-6, precision=precision)
    monkeypatch.setenv("CUPHU_LAPLACE_LPR", "64")
    res = {}
    for merge in ("0", "2"):
        monkeypatch.setenv("CUPHU_LAPLACE_MERGE", merge)
        res[merge] = cuphu.unwrap_batch([s.igram for s in scenes], [s.corr for s in scenes], 10.0, "defo",
                                        laplace_options=opts, return_stats=True)
    for i in range(len(scenes)):
        np.testing.assert_array_equal(res["0"][0][i], res["2"][0][i], err_msg=str(dims[i]))
        np.testing.assert_array_equal(res["0"][1][i], res["2"][1][i], err_msg=str(dims[i]))
        a, b = res["0"][2][i], res["2"][2][i]
        assert (a["status"], a["iters"], a["pass_status"]) == (b["status"], b["iters"], b["pass_status"])
        np.testing.assert_allclose(a["relres"], b["relres"], rtol=1e-9)


@pytest.mark.parametrize("solver", ["fused", "batched", "plain"])
def test_degenerate_shapes(gpu_ready, solver):
    """1x1, single rows/columns and 2x2 instances, alone and in one batch
    with a normal one: finite, congruent with the wrapped phase, and equal
    to the FP64 oracle's ambiguities."""
    shapes = [(1, 1), (1, 9), (9, 1), (2, 2), (3, 300), (300, 3)]
    rng = np.random.default_rng(17)
    scenes = []
    for nr, nc in shapes:
        ph = rng.uniform(-3, 3, (nr, nc))
        ig = np.exp(1j * np.cumsum(np.cumsum(ph * 0.3, 0), 1)).astype(np.complex64)
        corr = rng.uniform(0.3, 0.9, (nr, nc)).astype(np.float32)
        scenes.append((ig, corr))
    big = synth.make_scene(64, 80, nbumps=1, npatches=0, nlooks=5.0, seed=2)
    scenes.append((big.igram, big.corr))
    opts = dict(n_irls=2, max_iter=2000, tol=1e-6, solver=solver)
    unw, cc = cuphu.unwrap_batch([s[0] for s in scenes], [s[1] for s in scenes], 5.0, "defo", laplace_options=opts)
    for (ig, corr), u, c in zip(scenes, unw, cc):
        ph = O.wrapped_phase(ig, correctly_rounded=True)
        assert np.isfinite(u).all()
        k = (u.astype(np.float64) - ph) / TWOPI
        assert np.abs(k - np.rint(k)).max() < 1e-4
        sig = O.costs(ph, corr, 5.0, defo=True)[:, 1]
        uo, _ = O.pcg_irls_f64(sig, ph, n_irls=2, max_iter=2000, tol=1e-6)
        assert (np.rint((u - ph) / TWOPI) == np.rint((uo - ph) / TWOPI)).mean() >= 0.99
        assert c.shape == u.shape
    # each alone equals its batch result
    for i, (ig, corr) in enumerate(scenes[:-1]):
        u1, c1 = cuphu.unwrap_batch([ig], [corr], 5.0, "defo", laplace_options=opts)
        np.testing.assert_array_equal(u1[0], unw[i])


def test_many_small_instances(gpu_ready):
    """3000 interferograms of 4..96 px per side in one launch sequence: every
    instance finite and congruent, a sample bit-identical to its solo solve
    and agreeing with the FP64 oracle."""
    rng = np.random.default_rng(99)
    n = 3000
    igs, cos, shapes = [], [], []
    for i in range(n):
        nr, nc = int(rng.integers(4, 97)), int(rng.integers(4, 97))
        r, c = np.mgrid[0:nr, 0:nc]
        ph = 0.3 * r - 0.2 * c + rng.normal(0, 0.3, (nr, nc))
        igs.append(np.exp(1j * ph).astype(np.complex64))
        cos.append(rng.uniform(0.2, 0.95, (nr, nc)).astype(np.float32))
        shapes.append((nr, nc))
    opts = dict(n_irls=1, max_iter=200, tol=1e-6)
    unw, cc, st = cuphu.unwrap_batch(igs, cos, 5.0, "defo", laplace_options=opts, return_stats=True)
    assert len(unw) == n
    for u, sh in zip(unw, shapes):
        assert u.shape == sh and np.isfinite(u).all()
    for i in rng.choice(n, 12, replace=False):
        u1, c1 = cuphu.unwrap_batch([igs[i]], [cos[i]], 5.0, "defo", laplace_options=opts)
        np.testing.assert_array_equal(unw[i], u1[0])
        np.testing.assert_array_equal(cc[i], c1[0])
        ph = O.wrapped_phase(igs[i], correctly_rounded=True)
        k = (unw[i].astype(np.float64) - ph) / TWOPI
        assert np.abs(k - np.rint(k)).max() < 1e-4
        sig = O.costs(ph, cos[i], 5.0, defo=True)[:, 1]
        uo, _ = O.pcg_irls_f64(sig, ph, **opts)
        assert (np.rint((unw[i] - ph) / TWOPI) == np.rint((uo - ph) / TWOPI)).mean() >= 0.99

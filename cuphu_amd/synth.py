"""Seeded synthetic interferograms for tests and benchmarks (numpy only).

Scene model (one generator for every workload):

* truth phase: a range ramp ``ramp * col`` (default 2 rad / 100 px) plus K
  Gaussian bumps with amplitude ~ U(-40, 40) rad, sigma ~ U(100, 300) px and
  centres uniform over the scene;
* coherence: a smooth Gaussian-filtered N(0, 1) field (sigma 32 px,
  filtered on a grid coarsened 8x and bilinearly upsampled), min-max scaled
  to [0.2, 0.95];
* P decorrelated 300x300 patches at seeded positions with coherence 0.05
  and uniformly random truth phase;
* noise consistent with coherence g and L looks: with unit circular
  Gaussian n1, n2 per look, s1 = n1 and s2 = g n1 + sqrt(1-g^2) n2, the
  L-look product is sum s1 conj(s2) = g G + sqrt((1-g^2) G) Z with
  G = sum |n1|^2 ~ Gamma(L, 1) and Z ~ CN(0, 1) (exact in distribution),
  so igram = exp(i truth) * (g G + sqrt((1-g^2) G) Z) / L.

Outputs: igram complex64, corr float32 (= g), truth float64.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np

__all__ = ["Scene", "make_scene", "ramp_256", "config_scene", "batch_instances", "checksum"]


@dataclass
class Scene:
    igram: np.ndarray   # complex64
    corr: np.ndarray    # float32
    truth: np.ndarray   # float64
    nlooks: float
    seed: int


def _coherence(nrow: int, ncol: int, rng: np.random.Generator, sigma: float = 32.0,
               lo: float = 0.2, hi: float = 0.95) -> np.ndarray:
    from scipy import ndimage

    f = 8
    h, w = nrow // f + 3, ncol // f + 3
    coarse = ndimage.gaussian_filter(rng.standard_normal((h, w)), sigma / f, mode="reflect")
    fine = ndimage.zoom(coarse, f, order=1)[:nrow, :ncol]
    mn, mx = float(fine.min()), float(fine.max())
    return lo + (hi - lo) * (fine - mn) / max(mx - mn, 1e-30)


def _truth(nrow: int, ncol: int, nbumps: int, rng: np.random.Generator, ramp: float) -> np.ndarray:
    rr = np.arange(nrow, dtype=np.float64)
    cc = np.arange(ncol, dtype=np.float64)
    truth = np.empty((nrow, ncol), dtype=np.float64)
    truth[:] = ramp * cc[None, :]
    for _ in range(nbumps):
        amp = rng.uniform(-40.0, 40.0)
        sig = rng.uniform(100.0, 300.0)
        r0 = rng.uniform(0.0, nrow)
        c0 = rng.uniform(0.0, ncol)
        gr = np.exp(-((rr - r0) ** 2) / (2.0 * sig * sig))
        gc = np.exp(-((cc - c0) ** 2) / (2.0 * sig * sig))
        truth += amp * gr[:, None] * gc[None, :]
    return truth


def make_scene(nrow: int, ncol: int, *, nbumps: int, npatches: int, nlooks: float, seed: int,
               ramp: float = 0.02, patch: int = 300, patch_corr: float = 0.05) -> Scene:
    """Generate one synthetic scene (see module docstring)."""
    rng = np.random.default_rng(seed)
    truth = _truth(nrow, ncol, nbumps, rng, ramp)
    gam = _coherence(nrow, ncol, rng)
    ph, pw = min(patch, nrow), min(patch, ncol)
    for _ in range(npatches):
        r0 = int(rng.integers(0, nrow - ph + 1))
        c0 = int(rng.integers(0, ncol - pw + 1))
        gam[r0:r0 + ph, c0:c0 + pw] = patch_corr
        truth[r0:r0 + ph, c0:c0 + pw] = rng.uniform(-np.pi, np.pi, size=(ph, pw))
    L = float(nlooks)
    igram = np.empty((nrow, ncol), dtype=np.complex64)
    step = max(1, (1 << 22) // max(ncol, 1))
    for a in range(0, nrow, step):
        b = min(nrow, a + step)
        g = gam[a:b]
        G = rng.gamma(L, 1.0, size=g.shape)
        Z = (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)) * np.sqrt(0.5)
        prod = g * G + np.sqrt((1.0 - g * g) * G) * Z
        igram[a:b] = (np.exp(1j * truth[a:b]) * prod / L).astype(np.complex64)
    return Scene(igram=igram, corr=gam.astype(np.float32), truth=truth, nlooks=L, seed=seed)


def ramp_256() -> Scene:
    """The 256x256 ramp exp(i*pi*(x+y)), x (columns) and y (rows) on
    linspace(-3, 3, 256), corr = 1, nlooks = 1.  Closed-form unwrapped
    phase: pi*(x+y) + 6*pi."""
    x = np.linspace(-3.0, 3.0, 256)
    X, Y = np.meshgrid(x, x)
    truth = np.pi * (X + Y) + 6.0 * np.pi
    igram = np.exp(1j * np.pi * (X + Y)).astype(np.complex64)
    corr = np.ones((256, 256), dtype=np.float32)
    return Scene(igram=igram, corr=corr, truth=truth, nlooks=1.0, seed=0)


# Named workloads: (nrow, ncol, nbumps, npatches, nlooks, seed)
_CONFIGS = {
    "scene2048": (2048, 2048, 5, 1, 10.0, 2),
    "scene8192": (8192, 8192, 20, 4, 10.0, 3),
    "scene1024": (1024, 1024, 5, 1, 10.0, 5),
}


def config_scene(name: str, scale: float = 1.0) -> Scene:
    """One of the named scenes; ``scale`` < 1 shrinks it (tests)."""
    nrow, ncol, k, p, L, seed = _CONFIGS[name]
    nrow, ncol = max(8, int(nrow * scale)), max(8, int(ncol * scale))
    return make_scene(nrow, ncol, nbumps=k, npatches=p, nlooks=L, seed=seed)


def batch_instances(n: int = 64, seed: int = 4, scale: float = 1.0,
                    rows: tuple[int, int] = (512, 3000), cols: tuple[int, int] = (512, 1800),
                    first: int | None = None) -> list[Scene]:
    """n independent interferograms: nrow ~ U{rows}, ncol ~ U{cols},
    nlooks ~ U{1, 5, 10, 20} drawn from default_rng(seed); instance i is
    generated with seed + 1 + i (K = 5 bumps, P = 1 patch).  ``first``
    generates only the first instances of that same list."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n if first is None else min(n, first)):
        nr = int(rng.integers(rows[0], rows[1] + 1))
        nc = int(rng.integers(cols[0], cols[1] + 1))
        L = float(rng.choice([1, 5, 10, 20]))
        nr, nc = max(8, int(nr * scale)), max(8, int(nc * scale))
        out.append(make_scene(nr, nc, nbumps=5, npatches=1, nlooks=L, seed=seed + 1 + i))
    return out


def checksum(*arrays: np.ndarray) -> str:
    """sha256 over the raw bytes of the arrays (generator freeze check)."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()

"""Writes the golden fixture tests/golden/scene64x80.npz.

Inputs: one small seeded scene from the synthetic generator (a ramp of
0.25 rad/px, two Gaussian bumps, two 20 x 20 patches at coherence 0.02,
below the connected components' validity floor, noise at nlooks 5; a
64 x 80 complex64 interferogram and float32 coherence).  Expected outputs, from the
CPU oracle (oracle/, the C restatement of the reference path):

* phi: the wrapped phase in [0, 2 pi) (float32);
* costs_smooth / costs_defo: the SNAPHU arc costs, flat arc order
  (int16 [narcs, 2] offset/sigsq and [narcs, 4]);
* residues: the residue charges of the 2 x 2 loops (int8);
* unw: the FP64 Laplace PCG/IRLS solution (defo sigsq, 3 IRLS passes,
  400-iteration cap, tol 1e-6), anchored and congruent (float64);
* amb: its 2 pi ambiguity counts round((unw - phi) / 2 pi) (int32);
* labels: the connected components of unw (uint32).

Run from the repository root after building the oracle:
    python tests/golden/make_fixtures.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from cuphu_amd import synth  # noqa: E402

SCENE = dict(nrow=64, ncol=80, nbumps=2, npatches=2, nlooks=5.0, seed=7, patch=20, ramp=0.25, patch_corr=0.02)
SOLVE = dict(n_irls=3, max_iter=400, tol=1e-6)
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scene64x80.npz")


def make():
    s = synth.make_scene(SCENE["nrow"], SCENE["ncol"], nbumps=SCENE["nbumps"], npatches=SCENE["npatches"],
                         nlooks=SCENE["nlooks"], seed=SCENE["seed"], patch=SCENE["patch"], ramp=SCENE["ramp"],
                         patch_corr=SCENE["patch_corr"])
    phi = O.wrapped_phase(s.igram, correctly_rounded=True)
    costs_smooth = O.costs(phi, s.corr, s.nlooks, defo=False)
    costs_defo = O.costs(phi, s.corr, s.nlooks, defo=True)
    res = O.residues(phi)
    unw, _ = O.pcg_irls_f64(costs_defo[:, 1], phi, **SOLVE)
    amb = np.rint((unw - phi.astype(np.float64)) / (2 * np.pi)).astype(np.int32)
    labels = O.conncomp(unw.astype(np.float32), s.corr, None, 0.01, 32)
    return dict(igram=s.igram, corr=s.corr, nlooks=np.float64(s.nlooks), phi=phi, costs_smooth=costs_smooth,
                costs_defo=costs_defo, residues=res, unw=unw, amb=amb, labels=labels.astype(np.uint32),
                n_irls=np.int32(SOLVE["n_irls"]), max_iter=np.int32(SOLVE["max_iter"]), tol=np.float64(SOLVE["tol"]))


if __name__ == "__main__":
    O.build()
    np.savez_compressed(PATH, **make())
    print("wrote", PATH, os.path.getsize(PATH), "bytes")

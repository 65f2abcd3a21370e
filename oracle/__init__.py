"""CPU oracle for the Laplace unwrapping path -- TEST INFRASTRUCTURE ONLY.

numpy wrappers over the C restatement in cuphu_oracle.c (built by the
Makefile in this directory into oracle/_build/).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker; the product never calls it.

Every function below returns what the reference path computes for the
same inputs (see the file:line citations in cuphu_oracle.c); how each one
is pinned is listed in that file's header.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
# CUPHU_ORACLE_LIB selects another build of the oracle (the sanitizer lane)
LIB_PATH = os.environ.get("CUPHU_ORACLE_LIB") or os.path.join(_DIR, "_build", "libcuphu_oracle.so")


def build(force: bool = False) -> None:
    """Compile the oracle with gcc (make -C oracle)."""
    args = ["make", "-C", _DIR] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


class OCostParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "ncorrlooks", "rhosconst1", "rhosconst2", "cstd1", "cstd2", "cstd3", "defothreshfactor",
        "sigsqcorr", "costscale", "nshortcycle", "defomax")] + [("sigsqshortmin", C.c_long)]


class OLapOptions(C.Structure):
    _fields_ = [("n_irls", C.c_int), ("delta_max", C.c_float), ("delta_min", C.c_float),
                ("max_iter", C.c_int), ("tol", C.c_double), ("continuous", C.c_int)]


class OLapStats(C.Structure):
    _fields_ = [("npass", C.c_int), ("iters", C.c_int * 33), ("relres", C.c_double * 33),
                ("status", C.c_int * 33)]

    def as_dict(self) -> dict:
        n = self.npass
        return {"npass": n, "iters": list(self.iters[:n]), "relres": list(self.relres[:n]),
                "status": list(self.status[:n])}


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


_lib = _load()
_P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def narcs(nrow: int, ncol: int) -> int:
    return (nrow - 1) * ncol + nrow * (ncol - 1)


def set_threads(n: int) -> None:
    _lib.o_set_threads(C.c_int(n))


def cost_params(nlooks: float, **over) -> OCostParams:
    """SNAPHU documented defaults (the library's cuphu_default_params values)."""
    d = dict(ncorrlooks=nlooks, rhosconst1=1.3, rhosconst2=0.14, cstd1=0.4, cstd2=0.35, cstd3=0.06,
             defothreshfactor=1.2, sigsqcorr=0.05, costscale=100.0, nshortcycle=200.0, defomax=1.2,
             sigsqshortmin=1)
    d.update(over)
    return OCostParams(**d)


def wrapped_phase(igram: np.ndarray, correctly_rounded: bool = True) -> np.ndarray:
    ig = np.ascontiguousarray(igram, dtype=np.complex64)
    re, im = _f32(ig.real), _f32(ig.imag)
    out = np.empty(ig.shape, np.float32)
    _lib.o_wrapped_phase(_p(re), _p(im), C.c_size_t(ig.size), C.c_int(int(correctly_rounded)), _p(out))
    return out


def phase_diffs(phi: np.ndarray):
    phi = _f32(phi)
    nrow, ncol = phi.shape
    rng = np.empty((nrow, max(ncol - 1, 0)), np.float32)
    az = np.empty((max(nrow - 1, 0), ncol), np.float32)
    _lib.o_phase_diffs(_p(phi), C.c_int(nrow), C.c_int(ncol), _p(rng), _p(az))
    return rng, az


def boxcar2d(src: np.ndarray, kr: int, kc: int) -> np.ndarray:
    src = _f32(src)
    out = np.empty_like(src)
    _lib.o_boxcar2d(_p(src), C.c_int(src.shape[0]), C.c_int(src.shape[1]), C.c_int(kr), C.c_int(kc), _p(out))
    return out


def costs(phi, corr, nlooks, *, weights=None, kperp=7, kpar=7, defo=False, **over) -> np.ndarray:
    """Flat cost array [narcs, 2] (offset, sigsq) or [narcs, 4] for defo."""
    phi, corr = _f32(phi), _f32(corr)
    nrow, ncol = phi.shape
    ne = 4 if defo else 2
    out = np.empty((narcs(nrow, ncol), ne), np.int16)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int16)
    prm = cost_params(nlooks, **over)
    _lib.o_costs(_p(phi), _p(corr), _p(w), C.c_int(nrow), C.c_int(ncol), C.byref(prm), C.c_int(kperp),
                 C.c_int(kpar), C.c_int(int(defo)), _p(out))
    return out


def residues(phi) -> np.ndarray:
    phi = _f32(phi)
    nrow, ncol = phi.shape
    out = np.empty((nrow - 1, ncol - 1), np.int8)
    _lib.o_residues(_p(phi), C.c_int(nrow), C.c_int(ncol), _p(out))
    return out


def arc_weights(sigsq) -> np.ndarray:
    s = np.ascontiguousarray(sigsq, dtype=np.int16)
    out = np.empty(s.size, np.float32)
    _lib.o_arc_weights(_p(s), C.c_size_t(s.size), C.c_size_t(1), _p(out))
    return out


def assemble(w, phi):
    w, phi = _f32(w), _f32(phi)
    nrow, ncol = phi.shape
    b = np.empty_like(phi)
    diag = np.empty_like(phi)
    _lib.o_assemble(_p(w), _p(phi), C.c_int(nrow), C.c_int(ncol), _p(b), _p(diag))
    return b, diag


def matvec(w, x):
    w, x = _f32(w), _f32(x)
    y = np.empty_like(x)
    _lib.o_matvec(_p(w), _p(x), C.c_int(x.shape[0]), C.c_int(x.shape[1]), _p(y))
    return y


def reweight(sigsq, phi, u, delta: float):
    s = np.ascontiguousarray(sigsq, dtype=np.int16)
    phi, u = _f32(phi), _f32(u)
    out = np.empty(s.size, np.float32)
    _lib.o_reweight(_p(s), C.c_size_t(1), _p(phi), _p(u), C.c_float(delta), C.c_int(phi.shape[0]),
                    C.c_int(phi.shape[1]), _p(out))
    return out


def congruence(phi, u):
    phi = _f32(phi)
    u = _f32(u).copy()
    _lib.o_congruence(_p(phi), _p(u), C.c_size_t(u.size))
    return u


def _opts(n_irls=16, delta_max=0.5, delta_min=0.01, max_iter=1000, tol=1e-3, continuous=False) -> OLapOptions:
    """continuous=True skips the final congruence (anchored WLS solution)."""
    return OLapOptions(int(n_irls), float(delta_max), float(delta_min), int(max_iter), float(tol), int(continuous))


def pcg_irls_f32(sigsq, phi, *, u0=None, **opts):
    """FP32-vector / FP64-reduction restatement of the reference driver."""
    s = np.ascontiguousarray(sigsq, dtype=np.int16)
    phi = _f32(phi)
    u = np.empty_like(phi)
    st = OLapStats()
    o = _opts(**opts)
    u0a = None if u0 is None else _f32(u0)
    _lib.o_pcg_irls_f32(_p(s), C.c_size_t(1), _p(phi), C.c_int(phi.shape[0]), C.c_int(phi.shape[1]),
                        C.byref(o), _p(u0a), _p(u), C.byref(st))
    return u, st.as_dict()


def pcg_irls_f64(sigsq, phi, *, u0=None, **opts):
    """FP64 solve of the same weighted normal equations and schedule."""
    s = np.ascontiguousarray(sigsq, dtype=np.int16)
    phi = _f32(phi)
    u = np.empty(phi.shape, np.float64)
    st = OLapStats()
    o = _opts(**opts)
    u0a = None if u0 is None else _f32(u0)
    _lib.o_pcg_irls_f64(_p(s), C.c_size_t(1), _p(phi), C.c_int(phi.shape[0]), C.c_int(phi.shape[1]),
                        C.byref(o), _p(u0a), _p(u), C.byref(st))
    return u, st.as_dict()


def pcg_step_f32(w, dinv, gamma, beta, x, r, p, alpha=None):
    """One Jacobi-PCG iteration (direction update first).  Returns
    (x, r, p, q, pq, alpha, rz, status); status 2 = p.q <= 0."""
    w, dinv = _f32(w), _f32(dinv)
    x, r, p = (_f32(a).copy() for a in (x, r, p))
    q = np.zeros_like(x)
    pq, al, rz = C.c_double(), C.c_double(float("nan")), C.c_double(float("nan"))
    st = _lib.o_pcg_step_f32(_p(w), _p(dinv), C.c_int(x.shape[0]), C.c_int(x.shape[1]), C.c_double(gamma),
                             C.c_double(beta), C.c_double(float("nan") if alpha is None else alpha), _p(x), _p(r),
                             _p(p), _p(q), C.byref(pq), C.byref(al), C.byref(rz))
    return x, r, p, q, pq.value, al.value, rz.value, st


def pcg_fused_step_f32(w, dinv, alpha, beta, x, r, p, q, start=False, rt=None):
    """One fused PCG iteration.  Returns (x, r, p, q, sums[4])."""
    w, dinv = _f32(w), _f32(dinv)
    x, r, p, q = (_f32(a).copy() for a in (x, r, p, q))
    rta = None if rt is None else _f32(rt)
    sums = np.zeros(4, np.float64)
    _lib.o_pcg_fused_step_f32(_p(w), _p(dinv), C.c_int(x.shape[0]), C.c_int(x.shape[1]), C.c_double(alpha),
                              C.c_double(beta), C.c_int(1 if start else 0), _p(rta), _p(x), _p(r), _p(p), _p(q),
                              _p(sums))
    return x, r, p, q, sums


def conncomp(unw, corr, mask=None, minfrac=0.01, maxncomps=32, poscost=None, negcost=None,
             conncompthresh=300) -> np.ndarray:
    """Connected components; with poscost/negcost (flat arc arrays), the MCF
    path's incremental-cost gate min(pos, neg) >= conncompthresh."""
    unw, corr = _f32(unw), _f32(corr)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    pc = None if poscost is None else np.ascontiguousarray(poscost, dtype=np.int16)
    nc = None if negcost is None else np.ascontiguousarray(negcost, dtype=np.int16)
    out = np.empty(unw.shape, np.uint32)
    _lib.o_conncomp_gated(_p(unw), _p(corr), _p(m), C.c_int(unw.shape[0]), C.c_int(unw.shape[1]),
                          C.c_double(minfrac), C.c_long(maxncomps), _p(pc), _p(nc), C.c_long(conncompthresh),
                          _p(out))
    return out


def incrcost_smooth(costs, flows, nshortcycle=200, nflow=1):
    """Incremental costs of smooth-cost arcs at the given flows.  costs:
    int16 [narcs, 2] (offset, sigsq); flows int16 [narcs].  Returns
    (poscost, negcost, early_exit)."""
    cs = np.ascontiguousarray(costs, dtype=np.int16)
    fl = np.ascontiguousarray(flows, dtype=np.int16)
    n = fl.size
    pc = np.empty(n, np.int16)
    nc = np.empty(n, np.int16)
    ee = _lib.o_incrcost_smooth(_p(cs), _p(fl), C.c_size_t(n), C.c_long(nshortcycle), C.c_int(nflow), _p(pc), _p(nc))
    return pc, nc, bool(ee)


def integrate_phase(phi, hflows, vflows) -> np.ndarray:
    phi = _f32(phi)
    h = np.ascontiguousarray(hflows, dtype=np.int16)
    v = np.ascontiguousarray(vflows, dtype=np.int16)
    out = np.empty_like(phi)
    _lib.o_integrate_phase(_p(phi), _p(h), _p(v), C.c_int(phi.shape[0]), C.c_int(phi.shape[1]), _p(out))
    return out


def tile_axis(n: int, tile_size: int, overlap: int, ntiles: int = 1):
    f = (C.c_int * 4096)()
    s = (C.c_int * 4096)()
    k = _lib.o_tile_axis(C.c_int(n), C.c_int(tile_size), C.c_int(overlap), C.c_int(ntiles), f, s, C.c_int(4096))
    if k < 0:
        raise ValueError("invalid tiling")
    return list(f[:k]), list(s[:k])


def stitch(nrow, ncol, rf, rs, cf, cs, rovl, covl, tiles_unw, tiles_cc, tiles_corr, tiles_conf=None,
           laplace=True):
    """Stitch solved tiles (lists in row-major tile order)."""
    ntr, ntc = len(rf), len(cf)
    nt = ntr * ntc
    iv = lambda v: (C.c_int * len(v))(*v)  # noqa: E731
    keep = []

    def ptrs(arrs, dt):
        out = (C.c_void_p * nt)()
        for i, a in enumerate(arrs):
            if a is None:
                out[i] = None
                continue
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            out[i] = a.ctypes.data
        return out

    pu, pc, pr = ptrs(tiles_unw, np.float32), ptrs(tiles_cc, np.uint32), ptrs(tiles_corr, np.float32)
    pf = ptrs(tiles_conf, np.uint8) if tiles_conf is not None else None
    unw = np.empty((nrow, ncol), np.float32)
    cc = np.empty((nrow, ncol), np.uint32)
    tk = (C.c_double * nt)()
    _lib.o_stitch(C.c_int(nrow), C.c_int(ncol), C.c_int(ntr), C.c_int(ntc), iv(rf), iv(rs), iv(cf), iv(cs),
                  C.c_int(rovl), C.c_int(covl), C.c_int(int(laplace)), pu, pc, pr, pf, _p(unw), _p(cc), tk)
    return unw, cc, np.array(tk[:])


def fix_cycle_spikes(unw, mask=None) -> np.ndarray:
    """Isolated whole-cycle row/column spike correction (a new array)."""
    out = np.array(unw, dtype=np.float32, order="C", copy=True)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    _lib.o_fix_cycle_spikes(_p(out), _p(m), C.c_int(out.shape[0]), C.c_int(out.shape[1]))
    return out


def neighbor_feedback(unw, mask, ntilerow, ntilecol, row_ovrlp, col_ovrlp, feather_px) -> np.ndarray:
    """Laplace tile-seam neighbour feedback (a new array); seams at the
    reference tile formula's tile starts + overlap/2."""
    out = np.array(unw, dtype=np.float32, order="C", copy=True)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    _lib.o_neighbor_feedback(_p(out), _p(m), C.c_int(out.shape[0]), C.c_int(out.shape[1]), C.c_int(ntilerow),
                             C.c_int(ntilecol), C.c_int(row_ovrlp), C.c_int(col_ovrlp), C.c_int(feather_px))
    return out


RAMP_TYPES = {None: 0, "linear": 1, "quadratic": 2, "linear_range": 3, "linear_azimuth": 4,
              "quadratic_range": 5, "quadratic_azimuth": 6}


def bridge(unw, mask=None, radius=500, min_num_pixel=14, erosion_size=2, max_boundary_samples=4096,
           ramp_type=None, ramp_max_num_sample=1000000):
    """Phase bridging; returns (unw, cc, nregions) -- copies; cc is all 0
    and unw unchanged when fewer than two regions remain."""
    out = np.array(unw, dtype=np.float32, order="C", copy=True)
    cc = np.zeros(out.shape, np.uint32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    nreg = _lib.o_bridge(_p(out), _p(cc), _p(m), C.c_int(out.shape[0]), C.c_int(out.shape[1]), C.c_int(radius),
                         C.c_int(min_num_pixel), C.c_int(erosion_size), C.c_int(max_boundary_samples),
                         C.c_int(RAMP_TYPES[ramp_type]), C.c_int(ramp_max_num_sample))
    return out, cc, nreg


def bridge_label(topo):
    topo = _f32(topo)
    lab = np.zeros(topo.shape, np.uint32)
    n = _lib.o_bridge_label(_p(topo), C.c_int(topo.shape[0]), C.c_int(topo.shape[1]), _p(lab))
    return lab, n


def bridge_erode_prune(labels, n, e, circular):
    lab = np.array(labels, dtype=np.uint32, order="C", copy=True)
    k = _lib.o_bridge_erode_prune(_p(lab), C.c_int(lab.shape[0]), C.c_int(lab.shape[1]), C.c_int(n), C.c_int(e),
                                  C.c_int(int(circular)))
    return lab, k


def bridge_boundary(labels):
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    b = np.zeros(lab.shape, np.uint8)
    _lib.o_bridge_boundary(_p(lab), C.c_int(lab.shape[0]), C.c_int(lab.shape[1]), _p(b))
    return b


def bridge_nn(labels, boundary, n, max_samples=4096):
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    b = np.ascontiguousarray(boundary, dtype=np.uint8)
    dm = np.empty((n + 1) * (n + 1), np.float32)
    ep = np.empty((n + 1) * (n + 1) * 4, np.int32)
    _lib.o_bridge_nn(_p(lab), _p(b), C.c_int(lab.shape[0]), C.c_int(lab.shape[1]), C.c_int(n),
                     C.c_int(max_samples), _p(dm), _p(ep))
    return dm.reshape(n + 1, n + 1), ep.reshape(n + 1, n + 1, 4)


def wls_solve_f64(w, phi, max_iter=10000, tol=1e-12):
    w = np.ascontiguousarray(w, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    u = np.empty(phi.shape, np.float64)
    it = C.c_int()
    _lib.o_wls_solve_f64(_p(w), _p(phi), C.c_int(phi.shape[0]), C.c_int(phi.shape[1]), C.c_int(max_iter),
                         C.c_double(tol), _p(u), C.byref(it))
    return u, it.value


def wls_adjoint_f64(w, phi, u, gu, max_iter=10000, tol=1e-12):
    w = np.ascontiguousarray(w, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    gu = np.ascontiguousarray(gu, dtype=np.float64)
    gw = np.empty(w.size, np.float64)
    lam = np.empty(phi.shape, np.float64)
    _lib.o_wls_adjoint_f64(_p(w), _p(phi), _p(u), _p(gu), C.c_int(phi.shape[0]), C.c_int(phi.shape[1]),
                           C.c_int(max_iter), C.c_double(tol), _p(gw), _p(lam))
    return gw, lam


def _sig(fn, res, args):
    f = getattr(_lib, fn)
    f.restype = res
    f.argtypes = args


for _n in ("o_set_threads", "o_wrapped_phase", "o_phase_diffs", "o_boxcar2d", "o_costs", "o_residues",
           "o_arc_weights", "o_assemble", "o_matvec", "o_reweight", "o_congruence", "o_pcg_irls_f32",
           "o_pcg_irls_f64", "o_pcg_step_f32", "o_pcg_fused_step_f32", "o_conncomp", "o_conncomp_gated", "o_incrcost_smooth", "o_integrate_phase", "o_tile_axis", "o_stitch",
           "o_wls_solve_f64", "o_wls_adjoint_f64", "o_fix_cycle_spikes", "o_neighbor_feedback", "o_bridge",
           "o_bridge_label", "o_bridge_prune_small", "o_bridge_erode_prune", "o_bridge_boundary", "o_bridge_nn"):
    getattr(_lib, _n).restype = C.c_int if _n not in ("o_set_threads", "o_wrapped_phase", "o_phase_diffs",
                                                       "o_costs", "o_residues", "o_arc_weights",
                                                       "o_assemble", "o_matvec", "o_reweight",
                                                       "o_congruence", "o_pcg_fused_step_f32",
                                                       "o_integrate_phase", "o_fix_cycle_spikes",
                                                       "o_neighbor_feedback", "o_bridge_boundary",
                                                       "o_bridge_nn") else None

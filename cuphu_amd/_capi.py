"""ctypes binding of the C ABI in ``include/cuphu.h`` (libcuphu_hip.so).

The structs below mirror the C declarations field for field; a CPU test
compiles the header and compares ``sizeof``/``offsetof`` against them.
The library is loaded from this package directory only; a missing library
raises ImportError (there is no CPU fallback).
"""

from __future__ import annotations

import ctypes as C
import os

__all__ = [
    "lib", "CuPhuParams", "CuPhuTileParams", "CuPhuBridgeParams", "CuPhuResult",
    "CuPhuLaplaceOptions", "CuPhuSolveStats", "CuPhuInstance", "check", "LIB_PATH",
    "SOLVER_PLAIN", "SOLVER_BATCHED", "SOLVER_FUSED", "COST", "INIT", "EXPORTED",
]

# CUPHU_LIB selects another build of the library (the CPU sanitizer lane's
# build with an instrumented host stitch)
LIB_PATH = os.environ.get("CUPHU_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcuphu_hip.so")

COST = {"smooth": 0, "defo": 1, "topo": 2}
INIT = {"mst": 0, "mcf": 1, "laplace": 2}
SOLVER_PLAIN, SOLVER_BATCHED, SOLVER_FUSED = 0, 1, 2
MAX_IRLS = 32

OK, ERR_NOMEM, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_INTERNAL = 0, -1, -2, -3, -4, -5

_d, _i, _l = C.c_double, C.c_int, C.c_long


class CuPhuParams(C.Structure):
    _fields_ = [
        ("nlooks", _d), ("ncorrlooks", _d), ("defothreshfactor", _d), ("rhosconst1", _d),
        ("rhosconst2", _d), ("cstd1", _d), ("cstd2", _d), ("cstd3", _d), ("sigsqcorr", _d),
        ("costscale", _d), ("nshortcycle", _d), ("sigsqshortmin", _l),
        ("orbitradius", _d), ("altitude", _d), ("earthradius", _d), ("baseline", _d),
        ("baselineangle", _d), ("bperp", _d), ("transmitmode", _i), ("nlooksrange", _l),
        ("nlooksaz", _l), ("nlooksother", _l), ("ncorrlooksrange", _l), ("ncorrlooksaz", _l),
        ("nearrange", _d), ("dr", _d), ("da", _d), ("rangeres", _d), ("azres", _d), ("lambda_", _d),
        ("kds", _d), ("specularexp", _d), ("dzrcritfactor", _d), ("shadow", _i), ("dzeimin", _d),
        ("laywidth", _l), ("layminei", _d), ("sloperatiofactor", _d), ("sigsqei", _d),
        ("dzlaypeak", _d), ("azdzfactor", _d), ("dzeifactor", _d), ("dzeiweight", _d),
        ("dzlayfactor", _d), ("layconst", _d), ("layfalloffconst", _d), ("sigsqlayfactor", _d),
        ("defoazdzfactor", _d), ("defomax", _d), ("defolayconst", _d),
        ("kperpdpsi", _i), ("kpardpsi", _i), ("p", _d), ("maxcost", _d), ("costscaleambight", _d),
        ("dnomincangle", _d), ("initdzr", _d), ("initdzstep", _d), ("threshold", _d),
        ("initmaxflow", _l), ("arcmaxflowconst", _l), ("maxflow", _l), ("nshortcycleL", _l),
        ("cs2scalefactor", _l),
        ("minconncompfrac", _d), ("conncompthresh", _l), ("maxncomps", _l),
    ]


class CuPhuTileParams(C.Structure):
    _fields_ = [(n, _i) for n in (
        "ntilerow", "ntilecol", "rowovrlp", "colovrlp", "tilecostthresh", "minregionsize", "nproc",
        "ngpustreams", "single_tile_reoptimize", "laplace_neighbor_feedback",
        "laplace_neighbor_feedback_feather", "fix_cycle_spikes")]


class CuPhuBridgeParams(C.Structure):
    _fields_ = [(n, _i) for n in (
        "enabled", "radius", "min_num_pixel", "erosion_size", "max_boundary_samples", "ramp_type",
        "ramp_max_num_sample")]


class CuPhuResult(C.Structure):
    _fields_ = [("unw", C.POINTER(C.c_float)), ("conncomp", C.POINTER(C.c_uint32)),
                ("nrow", _i), ("ncol", _i)]


class CuPhuLaplaceOptions(C.Structure):
    _fields_ = [("n_irls", _i), ("delta_max", C.c_float), ("delta_min", C.c_float), ("max_iter", _i),
                ("tol", _d), ("solver", _i), ("precision", _i), ("rr_interval", _i),
                ("use_graph", _i), ("group_bytes_mb", _i), ("verbose", _i),
                ("autotune", _i), ("stitch_integer", _i), ("merge_labels", _i)]


class CuPhuSolveStats(C.Structure):
    _fields_ = [("status", _i), ("npass", _i), ("iters", _i * (MAX_IRLS + 1)),
                ("relres", _d * (MAX_IRLS + 1)), ("pass_status", _i * (MAX_IRLS + 1))]

    def as_dict(self) -> dict:
        n = self.npass
        return {"status": self.status, "npass": n, "iters": list(self.iters[:n]),
                "relres": list(self.relres[:n]), "pass_status": list(self.pass_status[:n])}


class CuPhuKernelProfile(C.Structure):
    _fields_ = [("samples", C.c_longlong), ("px_sum", _d), ("pq_ms_sum", _d), ("pq_bytes_sum", _d),
                ("xr_ms_sum", _d), ("xr_bytes_sum", _d), ("it_ms_sum", _d), ("it_bytes_sum", _d),
                ("iterations", C.c_longlong)]


class CuPhuInstance(C.Structure):
    _fields_ = [("nrow", _i), ("ncol", _i), ("d_igram", C.c_void_p), ("d_phase", C.c_void_p),
                ("d_corr", C.c_void_p), ("d_mask", C.c_void_p), ("d_u0", C.c_void_p),
                ("nlooks", _d), ("d_unw", C.c_void_p), ("d_conncomp", C.c_void_p)]


_P = C.c_void_p
_PF = C.POINTER(C.c_float)
_PS = C.POINTER(C.c_short)

# name -> (restype, argtypes)
_SIGS = {
    "cuphu_default_params": (None, [C.POINTER(CuPhuParams)]),
    "cuphu_default_tile_params": (None, [C.POINTER(CuPhuTileParams)]),
    "cuphu_default_bridge_params": (None, [C.POINTER(CuPhuBridgeParams)]),
    "cuphu_default_laplace_options": (None, [C.POINTER(CuPhuLaplaceOptions)]),
    "cuphu_last_error": (C.c_char_p, []),
    "cuphu_gpu_count": (_i, []),
    "cuphu_gpu_name": (_i, [_i, C.c_char_p, _i]),
    "cuphu_gpu_info": (_i, [_i, C.POINTER(C.c_longlong), C.POINTER(_i), C.c_char_p, _i]),
    "cuphu_unwrap": (_i, [_P, _P, _P, _P, _P, _i, _i, _i, _i, C.POINTER(CuPhuParams),
                          C.POINTER(CuPhuTileParams), C.POINTER(CuPhuBridgeParams), _P, _i,
                          C.POINTER(CuPhuResult)]),
    "cuphu_unwrap_ex": (_i, [_P, _P, _P, _P, _P, _i, _i, _i, _i, C.POINTER(CuPhuParams),
                             C.POINTER(CuPhuTileParams), C.POINTER(CuPhuBridgeParams), _P, _i,
                             C.POINTER(CuPhuLaplaceOptions), C.POINTER(CuPhuSolveStats),
                             C.POINTER(CuPhuResult)]),
    "cuphu_build_costs_gpu": (_i, [_P, _P, _P, _P, _P, _i, _i, _i, C.POINTER(CuPhuParams), _i,
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "cuphu_integrate_phase_gpu": (_i, [_P, _P, _P, _i, _i, _i, _P]),
    "cuphu_conncomp_gpu": (_i, [_P, _P, _P, _i, _i, _P, _P, C.POINTER(CuPhuParams), _i, _P]),
    "cuphu_regrow_conncomp": (_i, [_P, _P, _P, _i, _i, _d, _d, _l, _i, _P]),
    "cuphu_wrap_phase": (_i, [_P, _i, _P]),
    "cuphu_phase_diffs": (_i, [_P, _P, _P, _i, _i, _P]),
    "cuphu_boxcar2d": (_i, [_P, _P, _i, _i, _i, _i, _P]),
    "cuphu_build_smooth_costs_gpu": (_i, [_P, _P, _P, _P, _i, _i, C.POINTER(CuPhuParams), _i, _i,
                                          C.POINTER(C.c_void_p), _P]),
    "cuphu_build_defo_costs_gpu": (_i, [_P, _P, _P, _P, _i, _i, C.POINTER(CuPhuParams), _i, _i,
                                        C.POINTER(C.c_void_p), _P]),
    "cuphu_laplace_unwrap_gpu": (_i, [_P, _P, _i, _i, _d, _i, C.c_float, _i, _P, _P]),
    "cuphu_laplace_unwrap_gpu_noalloc": (_i, [_P, _P, _i, _i, _d, _i, C.c_float, _i, _P, _P]),
    "cuphu_residues_gpu": (_i, [_P, _i, _i, _P, _P]),
    "cuphu_laplace_batch_gpu": (_i, [C.POINTER(CuPhuInstance), _i, _i, C.POINTER(CuPhuParams),
                                     C.POINTER(CuPhuLaplaceOptions), C.POINTER(CuPhuSolveStats), _P]),
    "cuphu_unwrap_scene_gpu": (_i, [_P, _P, _P, _P, _i, _i, _d, _i, C.POINTER(CuPhuParams), _i, _i,
                                    _i, _i, C.POINTER(CuPhuLaplaceOptions), _P, _P,
                                    C.POINTER(CuPhuSolveStats), C.POINTER(_i), _P]),
    "cuphu_tile_axis": (_i, [_i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i), _i]),
    "cuphu_stitch_tiles_host": (_i, [_i, _i, _i, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                     C.POINTER(_i), _i, _i, _i, C.POINTER(_P), C.POINTER(_P),
                                     C.POINTER(_P), C.POINTER(_P), _P, _P, C.POINTER(_d)]),
    "cuphu_wls_solve_gpu": (_i, [_P, _P, _i, _i, _i, _i, _d, _P, C.POINTER(_i), C.POINTER(_d), _P]),
    "cuphu_wls_adjoint_gpu": (_i, [_P, _P, _P, _P, _i, _i, _i, _i, _d, _P, _P, C.POINTER(_i),
                                   C.POINTER(_d), _P]),
    "cuphu_set_profiling": (_i, [_i]),
    "cuphu_get_profile": (_i, [C.POINTER(CuPhuKernelProfile)]),
    "cuphu_test_ingest": (_i, [_P, _P, _P, _i, _i, C.POINTER(CuPhuParams), _P, _P]),
    "cuphu_test_laplace_ops": (_i, [_P, _P, _P, _i, _i, C.c_float, _P, _P, _P, _P, _P, _P, _P]),
    "cuphu_test_pcg_step": (_i, [_P, _P, _i, _i, _d, _d, _i, _P, _P, _P, _P, C.POINTER(_d), C.POINTER(_d),
                                 C.POINTER(_d)]),
    "cuphu_test_pcg_fused_step": (_i, [_P, _P, _i, _i, _d, _d, _i, _P, _i, _i, _P, _P, _P, _P, _P]),
    "cuphu_fix_cycle_spikes_gpu": (_i, [_P, _P, _i, _i, _P]),
    "cuphu_neighbor_feedback_gpu": (_i, [_P, _P, _i, _i, _i, _i, _i, _i, _i, _P]),
    "cuphu_fix_cycle_spikes_test": (_i, [_P, _P, _i, _i]),
    "cuphu_laplace_neighbor_feedback_test": (_i, [_P, _P, _i, _i, _i, _i, _i, _i, _i]),
    "cuphu_bridge_gpu": (_i, [_P, _P, _P, _i, _i, C.POINTER(CuPhuBridgeParams), _P]),
    "cuphu_bridge_apply_test": (_i, [_P, _P, _i, _i, C.POINTER(CuPhuBridgeParams), _P, _i]),
    "cuphu_incrcost_gpu": (_i, [_P, _P, _i, _l, _i, _P, _P, C.POINTER(_i), _P]),
    "cuphu_set_flow_solver": (_i, [_P, _P]),
}

EXPORTED = tuple(_SIGS)


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libcuphu_hip.so not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    from cuphu_amd._runtime import ensure_single_runtime

    ensure_single_runtime()
    handle = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


lib = _load()


class CuPhuError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, where: str):
        self.code = code
        msg = lib.cuphu_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {code}: {msg}")


# int (*)(const void *costs, int cost_mode, const float *phase, int nrow, int ncol, int init_method,
#        short *flows_out, void *user)
FLOW_SOLVER_FN = C.CFUNCTYPE(_i, C.c_void_p, _i, C.c_void_p, _i, _i, _i, C.c_void_p, C.c_void_p)


def check(code: int, where: str) -> int:
    if code < 0:
        raise CuPhuError(code, where)
    return code


def default_params() -> CuPhuParams:
    p = CuPhuParams()
    lib.cuphu_default_params(C.byref(p))
    return p


def default_options(**over) -> CuPhuLaplaceOptions:
    o = CuPhuLaplaceOptions()
    lib.cuphu_default_laplace_options(C.byref(o))
    for k, v in over.items():
        if k == "solver" and isinstance(v, str):
            v = {"plain": SOLVER_PLAIN, "batched": SOLVER_BATCHED, "fused": SOLVER_FUSED}[v]
        if not hasattr(o, k):
            raise ValueError(f"unknown laplace option {k!r}")
        setattr(o, k, v)
    return o


def ptr(a) -> int | None:
    """Address of a numpy array / torch tensor (None for None)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data

#!/usr/bin/env python3
"""Benchmark of the Laplace PCG/IRLS unwrapping path on one MI355X.

Default workload (``scene8192``): a synthetic 8192x8192 interferogram (ramp
+ 20 Gaussian bumps, 4 decorrelated 300x300 patches, nlooks 10) unwrapped
by ``cuphu_unwrap_scene_gpu``: fixed 2048x2048 tiles with 256-px overlap
(25 tiles, ragged edge tiles), defo statistical cost, 5 IRLS re-solves
(6 PCG solves), tolerance 1e-6, the reference's 1000-iteration cap, all
tiles in one batched launch sequence, tile stitching and conncomp.  Inputs
are resident in HBM before the timed region; a step is one full unwrap.

Reported beside the headline Mpixel/s:
* roofline: the fused PCG iteration kernel k_it (x, r, p, q updated and
  q = L p formed in one pass; 40 B per active pixel) -- algorithmic bytes
  over its duration, timed live with HIP events on the solver's stream
  during the timed steps (with --solver batched: the direction+stencil
  kernel k_pq, 28 B/px, with the update kernel k_xr beside it);
* plain: the reference launch sequence (per-tile loop, 7 launches and 2
  blocking scalar reads per iteration) on the same scene, and the ratio;
* accuracy: 2-pi ambiguity agreement of the FP32 result with the FP64 mode
  of the same solver (same tiles and schedule), and unwrapped-phase error;
* cpu_baseline: the FP64 CPU oracle (same algorithm and schedule,
  OpenMP over the host cores this process may use, at most 16) on a
  1024x1024 crop of the scene.

Usage: python bench.py [--gpus 1] [--steps 2] [--warmup 1]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpixel/s unwrapped (Laplace PCG/IRLS) + stencil HBM GB/s; batched vs plain"
HBM_PEAK_GBPS = 8000.0
BYTES_PER_PX_IT = 40.0
BYTES_PER_PX_PQ = 28.0
BYTES_PER_PX_XR = 28.0
# HBM bytes per active pixel actually moved per launch at this workload:
# rocprofv3 FETCH_SIZE (x2.000, calibrated on a copy of known size) +
# WRITE_SIZE (x1.000) per launch / active pixels per launch
# (profiles/k_it_scene8192_hbm_traffic.md, profiles/k_pq_k_xr_scene8192_hbm_traffic.md).
# Re-measure after kernel changes.
TRAFFIC_PER_PX_PQ = 2770.6e6 / 84.93e6
TRAFFIC_PER_PX_IT = 41.73


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="scene8192", choices=["scene8192", "scene2048"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1024)
    ap.add_argument("--group-mb", type=int, default=0)
    # counter runs: fewer iterations (same launch shape), and a copy kernel of
    # known size to calibrate FETCH_SIZE / WRITE_SIZE
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--n-irls", type=int, default=5)
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--solver", default="fused", choices=["fused", "batched"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (tiles would be sharded across "
              "GPUs by a per-rank batch; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    from cuphu_amd import _capi, synth

    lib = _capi.lib
    if lib.cuphu_gpu_count() < 1 or not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    name = C.create_string_buffer(256)
    lib.cuphu_gpu_name(0, name, 256)

    if args.workload == "scene8192":
        tile, ovl = 2048, 256
    else:
        tile, ovl = 0, 0
    log(f"generating {args.workload}")
    t0 = time.perf_counter()
    scene = synth.config_scene(args.workload)
    log(f"generated in {time.perf_counter() - t0:.1f}s")
    nrow, ncol = scene.corr.shape
    npx = nrow * ncol
    dig = torch.from_numpy(scene.igram.view(np.float32).reshape(nrow, 2 * ncol)).to("cuda")
    dcorr = torch.from_numpy(scene.corr).to("cuda")
    dunw = torch.empty((nrow, ncol), dtype=torch.float32, device="cuda")
    dcc = torch.empty((nrow, ncol), dtype=torch.int32, device="cuda")
    prm = _capi.default_params()
    prm.nlooks = prm.ncorrlooks = scene.nlooks
    cost = _capi.COST["defo"]
    base_opts = dict(n_irls=args.n_irls, tol=1e-6, max_iter=args.max_iter, rr_interval=50,
                     group_bytes_mb=args.group_mb, solver=args.solver)
    if args.calib_copy:
        # 512 MiB read + 512 MiB written per launch (torch.neg: one elementwise pass)
        src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    torch.cuda.synchronize()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run(unw, cc, **over):
        o = _capi.default_options(**{**base_opts, **over})
        stats = (_capi.CuPhuSolveStats * 4096)()
        nt = C.c_int()
        rc = lib.cuphu_unwrap_scene_gpu(C.c_void_p(dig.data_ptr()), C.c_void_p(dcorr.data_ptr()), None, None,
                                        nrow, ncol, scene.nlooks, cost, C.byref(prm), tile, ovl, 1, 1, C.byref(o),
                                        C.c_void_p(unw.data_ptr()), C.c_void_p(cc.data_ptr()) if cc is not None else None,
                                        stats, C.byref(nt), stream)
        _capi.check(rc, "cuphu_unwrap_scene_gpu")
        return [stats[i].as_dict() for i in range(nt.value)]

    for w in range(args.warmup):
        log(f"warmup {w + 1}/{args.warmup}")
        st = run(dunw, dcc)
    torch.cuda.synchronize()
    lib.cuphu_set_profiling(1)
    log(f"timing {args.steps} step(s)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = run(dunw, dcc)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    prof = _capi.CuPhuKernelProfile()
    lib.cuphu_get_profile(C.byref(prof))
    lib.cuphu_set_profiling(0)
    ms_per_step = 1000.0 * t_total / args.steps
    mpx = npx * args.steps / t_total / 1e6
    unw32 = dunw.cpu().numpy()
    iters = [s["iters"] for s in st]
    tile_px = None
    log(f"batched: {mpx:.2f} Mpx/s ({ms_per_step:.0f} ms/step)")

    ns = max(prof.samples, 1)
    gbs = lambda b, ms: (b / 1e9) / (ms / 1e3) if ms > 0 else None
    if args.solver == "fused":
        kname, bpp, tpp = "k_it (fused PCG iteration: x, r, p, q updates + weighted 5-point stencil + 4 sums)", \
            BYTES_PER_PX_IT, TRAFFIC_PER_PX_IT
        k_ms, k_bytes = prof.it_ms_sum / ns, prof.it_bytes_sum / ns
        achieved = gbs(prof.it_bytes_sum, prof.it_ms_sum)
        extra = {}
    else:
        kname, bpp, tpp = "k_pq (PCG direction + weighted 5-point stencil + p.q)", BYTES_PER_PX_PQ, TRAFFIC_PER_PX_PQ
        k_ms, k_bytes = prof.pq_ms_sum / ns, prof.pq_bytes_sum / ns
        achieved = gbs(prof.pq_bytes_sum, prof.pq_ms_sum)
        xr_gbs = gbs(prof.xr_bytes_sum, prof.xr_ms_sum)
        extra = {"update_kernel": {"kernel": "k_xr (x, r update + r.z)", "bytes_per_px": BYTES_PER_PX_XR,
                                   "avg_launch_ms": round(prof.xr_ms_sum / ns, 4),
                                   "achieved": round(xr_gbs, 1) if xr_gbs else None,
                                   "frac": round(xr_gbs / HBM_PEAK_GBPS, 4) if xr_gbs else None}}
    traffic = round(k_bytes / bpp * tpp) if (tpp and args.workload == "scene8192") else None

    out = {
        "metric": METRIC,
        "value": round(mpx, 4),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ramp+bumps+coherence generator, cuphu_amd/synth.py)",
        "config": {
            "workload": f"{args.workload}: {nrow}x{ncol}, tiles {tile}/{ovl} overlap, defo cost, nlooks "
                        f"{scene.nlooks:g}, {args.n_irls} IRLS re-solves, tol 1e-6, max_iter {args.max_iter}, "
                        f"{args.solver} FP32/FP64-reduction PCG",
            "tiles": len(st), "n_irls": args.n_irls, "tol": 1e-6, "max_iter": args.max_iter, "rr_interval": 50,
            "solver": args.solver,
            "gpu": name.value.decode(),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4) if achieved else None,
            "traffic": traffic,
            "bytes_per_px": bpp,
            "avg_launch_ms": round(k_ms, 4),
            "avg_bytes_per_launch": k_bytes,
            "timed_launches": prof.samples,
            **extra,
        },
        "pcg_iters_per_tile": iters[:3] + (["..."] if len(iters) > 3 else []),
        "pcg_iters_total": int(sum(sum(i) for i in iters)),
    }

    if not args.no_plain:
        log("plain per-tile loop (reference launch sequence)")
        dunw_p = torch.empty_like(dunw)
        t0 = time.perf_counter()
        stp = run(dunw_p, None, solver="plain")
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain_mpx = npx / tp / 1e6
        up = dunw_p.cpu().numpy()
        phi = np.mod(np.angle(scene.igram.astype(np.complex128)), 2 * np.pi)
        a32 = np.rint((unw32 - phi) / (2 * np.pi))
        ap_ = np.rint((up - phi) / (2 * np.pi))
        out["plain"] = {"mpix_s": round(plain_mpx, 4), "seconds": round(tp, 2),
                        "batched_vs_plain": round(mpx / plain_mpx, 3),
                        "ambiguity_agreement_with_batched": float((a32 == ap_).mean()),
                        "pcg_iters_total": int(sum(sum(s["iters"]) for s in stp))}
        log(f"plain: {plain_mpx:.2f} Mpx/s, batched/plain = {mpx / plain_mpx:.2f}")
        del dunw_p

    if not args.no_fp64:
        log("FP64 reference mode (same tiles and schedule)")
        dunw64 = torch.empty_like(dunw)
        t0 = time.perf_counter()
        run(dunw64, None, precision=64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        u64 = dunw64.cpu().numpy()
        phi = np.mod(np.angle(scene.igram.astype(np.complex128)), 2 * np.pi)
        agree = float((np.rint((unw32 - phi) / (2 * np.pi)) == np.rint((u64 - phi) / (2 * np.pi))).mean())
        err = (unw32.astype(np.float64) - u64)
        out["accuracy_vs_fp64"] = {"ambiguity_agreement": agree, "max_abs_rad": float(np.abs(err).max()),
                                   "rms_rad": float(np.sqrt(np.mean(err * err))), "fp64_seconds": round(t64, 2)}
        log(f"agreement with FP64: {agree:.6f}")
        del dunw64

    if not args.no_cpu:
        import oracle as O

        n = min(args.cpu_sample, nrow, ncol)
        try:
            ncore = len(os.sched_getaffinity(0))
        except AttributeError:
            ncore = os.cpu_count() or 1
        ncore = max(1, min(ncore, 16))
        O.set_threads(ncore)
        ig = np.ascontiguousarray(scene.igram[:n, :n])
        co = np.ascontiguousarray(scene.corr[:n, :n])
        ph = O.wrapped_phase(ig)
        sig = O.costs(ph, co, scene.nlooks, defo=True)[:, 1]
        log(f"CPU oracle on {n}x{n} with {ncore} threads")
        t0 = time.perf_counter()
        _, cst = O.pcg_irls_f64(sig, ph, n_irls=args.n_irls, tol=1e-6, max_iter=args.max_iter)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(n * n / tc / 1e6, 5), "unit": "Mpixel/s", "cores": ncore, "kind": "port",
            "sample": f"FP64 C oracle (OpenMP) PCG/IRLS, same schedule ({args.n_irls + 1} solves, tol 1e-6, cap "
                      f"{args.max_iter}), on the "
                      f"{n}x{n} top-left crop of the scene; iterations {cst['iters']}; {tc:.1f} s; SNAPHU's CPU MCF "
                      f"is not available (source absent) and is not timed",
        }
        log(f"cpu: {n * n / tc / 1e6:.4f} Mpx/s")

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

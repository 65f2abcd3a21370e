"""``unwrap()``: the Python front door of the library.

Signature, defaults, argument meaning and error behaviour are those of the
reference's ``cuphu.unwrap`` (src/cuphu/_unwrap.py:167-451).  A call is
planned in small steps, each its own function:

* ``_plan_tiling``: tile counts and overlaps -- the caller's ``ntiles``, or
  the reference's tile-count formula toward ``target_tile_size``
  (``_auto_ntiles_by_target_size``);
* ``_plan_masks``: the mask the solver sees (the caller's mask grown by
  ``mask_pad_distance`` pixels of Manhattan distance) and where the seam
  feedback runs (inside the solve, or afterwards on the caller's mask when
  the solve mask was grown);
* the solve itself, through the pybind11 module ``_cuphu_ext`` on the gfx950
  library (non-finite inputs are zeroed there, on the device);
* ``_finish``: the seam feedback on the caller's mask, labels cleared
  outside it, the reference-pixel shift by whole cycles and the caller's
  output arrays.

``init='laplace'`` is the GPU path; ``'mcf'``/``'mst'`` need a network-flow
solver registered with ``cuphu_amd.set_flow_solver`` (SNAPHU's CPU solver
is not part of this library).

One keyword is added: ``laplace_options`` (dict; keys n_irls, delta_max,
delta_min, max_iter, tol, solver ('plain'|'batched'|'fused'), precision
(32|64), rr_interval, use_graph, group_bytes_mb, verbose, autotune, and for
tiled runs stitch_integer (whole-cycle tile offsets) and merge_labels
(connected components labelled on the stitched scene)).  Unset keys follow
the reference defaults and its CUPHU_* environment overrides.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from cuphu_amd._check import Raster, check_cost_mode, check_init_method, check_rasters
from cuphu_amd._ext import _cuphu_ext

__all__ = ["unwrap"]

TWO_PI = 2.0 * math.pi
# tile overlap when the caller gives none: 64 px for Laplace runs, none for
# the network-flow initialisers (reference _unwrap.py:28)
LAPLACE_DEFAULT_OVERLAP = 64


@dataclass(frozen=True)
class Tiling:
    rows: int
    cols: int
    row_overlap: int
    col_overlap: int


@dataclass(frozen=True)
class MaskPlan:
    solve: np.ndarray | None       # uint8 mask the solver sees
    original: np.ndarray | None    # the caller's mask (uint8)
    grown: bool                    # solve mask grown by mask_pad_distance
    feedback_in_solve: bool        # seam feedback inside the library call
    feedback_after: bool           # seam feedback after it, on the caller's mask


def _auto_ntiles_by_target_size(nrow: int, ncol: int, target_tile_size: int, row_ovrlp: int,
                                col_ovrlp: int) -> tuple[int, int]:
    """Tile counts whose tiles share one edge length near target_tile_size.

    The reference's formula (_unwrap.py:31-61): the longer axis is split into
    a whole number of steps of about target - overlap (the larger overlap),
    the resulting edge length is kept for both axes, and each axis gets the
    nearest whole number of such edges (at least one).
    """
    longest = max(nrow, ncol)
    step = max(1, target_tile_size - max(row_ovrlp, col_ovrlp))
    edge = longest / max(1, round(longest / step))
    return max(1, round(nrow / edge)), max(1, round(ncol / edge))


def _plan_tiling(shape: tuple[int, int], init: str, ntiles, tile_overlap, target_tile_size: int,
                 nproc: int) -> Tiling:
    if tile_overlap is None:
        tile_overlap = LAPLACE_DEFAULT_OVERLAP if init == "laplace" else 0
    rov, cov = (tile_overlap, tile_overlap) if np.ndim(tile_overlap) == 0 else tile_overlap
    rov, cov = int(rov), int(cov)
    if ntiles is not None:
        rows, cols = int(ntiles[0]), int(ntiles[1])
    elif init == "laplace" or nproc > 1:           # reference: auto-tiling unless one serial MCF tile
        rows, cols = _auto_ntiles_by_target_size(shape[0], shape[1], target_tile_size, rov, cov)
    else:
        rows, cols = 1, 1
    return Tiling(rows, cols, rov, cov)


def _dilate_mask(valid: np.ndarray, pixels: int) -> np.ndarray:
    """Pixels within Manhattan distance `pixels` of a valid pixel (the
    diamond the reference grows with 4-neighbour steps, _unwrap.py:64-84),
    from the exact city-block distance transform."""
    valid = np.asarray(valid, dtype=bool)
    if pixels <= 0 or not valid.any():
        return valid.copy()
    from scipy.ndimage import distance_transform_cdt

    return distance_transform_cdt(~valid, metric="taxicab") <= pixels


def _plan_masks(mask_u8: np.ndarray | None, pad: int, neighbor_feedback: bool,
                single_tile_reoptimize: bool) -> MaskPlan:
    """single_tile_reoptimize supersedes the seam feedback; with a grown
    solve mask the feedback runs afterwards on the caller's mask, so padded
    pixels never serve as seam samples (reference _unwrap.py:529-557)."""
    feedback = bool(neighbor_feedback) and not single_tile_reoptimize
    grown = mask_u8 is not None and pad > 0
    solve = _dilate_mask(mask_u8 != 0, int(pad)).astype(np.uint8) if grown else mask_u8
    return MaskPlan(solve=solve, original=mask_u8, grown=grown, feedback_in_solve=feedback and not grown,
                    feedback_after=feedback and grown)


def _wrapped_phase_at(z: complex) -> float:
    """Phase of one interferogram element in [0, 2 pi); a non-finite element
    is zeroed by the solve, so its phase is 0."""
    if not np.isfinite(z):
        return 0.0
    a = float(np.angle(z))
    return a + TWO_PI if a < 0.0 else a


def _finish(unw_out: np.ndarray, cc_out: np.ndarray, masks: MaskPlan, tiling: Tiling, feather: int,
            igram_c64: np.ndarray, reference_pixel, unw, conncomp):
    if masks.feedback_after:
        unw_out = _cuphu_ext._laplace_neighbor_feedback_test(
            np.ascontiguousarray(unw_out, dtype=np.float32), masks.original, tiling.rows, tiling.cols,
            tiling.row_overlap, tiling.col_overlap, int(feather))
    if masks.grown:
        # pixels admitted only by the padding are never reported valid (_unwrap.py:600-603)
        cc_out = np.where(masks.original != 0, cc_out, 0).astype(cc_out.dtype)
    if reference_pixel is not None:
        # whole-cycle shift: the reference pixel ends at its own wrapped phase (_unwrap.py:605-617)
        r, c = int(reference_pixel[0]), int(reference_pixel[1])
        cycles = round((float(unw_out[r, c]) - _wrapped_phase_at(igram_c64[r, c])) / TWO_PI)
        if cycles:
            unw_out = unw_out - np.float32(cycles * TWO_PI)
    if unw is not None:
        unw[...] = unw_out
        unw_out = unw
    if conncomp is not None:
        conncomp[...] = cc_out.astype(conncomp.dtype)
        cc_out = conncomp
    return unw_out, cc_out


def unwrap(
    igram,
    corr,
    nlooks,
    cost="smooth",
    init="mcf",
    *,
    mask=None,
    mask_pad_distance=0,
    mag=None,
    min_conncomp_frac=0.01,
    phase_grad_window=(7, 7),
    ntiles=None,
    tile_overlap=None,
    target_tile_size=1024,
    nproc=1,
    tile_cost_thresh=500,
    min_region_size=100,
    single_tile_reoptimize=False,
    laplace_neighbor_feedback=False,
    laplace_neighbor_feedback_feather=200,
    fix_cycle_spikes=False,
    bridge=False,
    bridge_radius=500,
    bridge_min_num_pixel=14,
    bridge_erosion_size=2,
    bridge_max_boundary_samples=4096,
    bridge_ramp_type=None,
    bridge_ramp_max_num_sample=1000000,
    reference_pixel=(0, 0),
    gpu_id=0,
    unw=None,
    conncomp=None,
    laplace_options=None,
):
    """Unwrap an interferogram on the GPU.

    Parameters and returns follow ``cuphu.unwrap`` (reference
    src/cuphu/_unwrap.py:201-451): *igram* complex 2-D, *corr* float 2-D,
    *nlooks* >= 1; returns ``(unw float32, conncomp uint32)`` (or the
    caller's *unw*/*conncomp* objects when given).
    """
    shape, arr = check_rasters([Raster("igram", igram, "complex"), Raster("corr", corr, "float"),
                                Raster("mask", mask, "mask"), Raster("mag", mag, "float")])
    check_cost_mode(cost)
    check_init_method(init)
    if nlooks < 1.0:
        raise ValueError(f"nlooks must be >= 1, got {nlooks}")
    if nproc < 1:
        nproc = os.cpu_count() or 1

    tiling = _plan_tiling(shape, init, ntiles, tile_overlap, target_tile_size, nproc)
    igram_c64 = np.ascontiguousarray(arr["igram"], dtype=np.complex64)
    mask_u8 = np.ascontiguousarray(arr["mask"], dtype=np.uint8) if "mask" in arr else None
    masks = _plan_masks(mask_u8, mask_pad_distance, laplace_neighbor_feedback, single_tile_reoptimize)

    unw_out, cc_out = _cuphu_ext.unwrap_arrays(
        igram_c64,
        np.ascontiguousarray(arr["corr"], dtype=np.float32),
        float(nlooks),
        cost=cost,
        init=init,
        mask=masks.solve,
        orig_mask=masks.original,
        mag=np.ascontiguousarray(arr["mag"], dtype=np.float32) if "mag" in arr else None,
        kperpdpsi=int(phase_grad_window[0]),
        kpardpsi=int(phase_grad_window[1]),
        min_conncomp_frac=float(min_conncomp_frac),
        ntilerow=tiling.rows,
        ntilecol=tiling.cols,
        tile_rowovrlp=tiling.row_overlap,
        tile_colovrlp=tiling.col_overlap,
        tilecostthresh=tile_cost_thresh,
        minregionsize=min_region_size,
        nproc=nproc,
        single_tile_reoptimize=bool(single_tile_reoptimize),
        laplace_neighbor_feedback=masks.feedback_in_solve,
        laplace_neighbor_feedback_feather=int(laplace_neighbor_feedback_feather),
        fix_cycle_spikes=bool(fix_cycle_spikes),
        bridge=bool(bridge),
        bridge_radius=int(bridge_radius),
        bridge_min_num_pixel=int(bridge_min_num_pixel),
        bridge_erosion_size=int(bridge_erosion_size),
        bridge_max_boundary_samples=int(bridge_max_boundary_samples),
        bridge_ramp_type=bridge_ramp_type,
        bridge_ramp_max_num_sample=int(bridge_ramp_max_num_sample),
        gpu_id=gpu_id,
        laplace_options=laplace_options,
    )
    return _finish(unw_out, cc_out, masks, tiling, laplace_neighbor_feedback_feather, igram_c64,
                   reference_pixel, unw, conncomp)

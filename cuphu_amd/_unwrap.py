"""``unwrap()``: the Python front door (reference: src/cuphu/_unwrap.py:167-629).

Same signature, defaults, validation and post-processing as the reference:
NaN scrub, auto-tiling toward ``target_tile_size``, mask dilation for
``mask_pad_distance``, the reference-pixel integer-cycle shift and writing
into caller-provided outputs, and the optional post-passes
``laplace_neighbor_feedback`` and ``fix_cycle_spikes``.  The solve itself goes
through the pybind11 module to the gfx950 library.  ``init='laplace'`` is
the GPU path;
``'mcf'``/``'mst'`` need a network-flow solver registered with
``cuphu_amd.set_flow_solver`` (SNAPHU's CPU solver is not part of this
library) and run on one tile; without one they raise.

One keyword is added: ``laplace_options`` (dict; keys n_irls, delta_max,
delta_min, max_iter, tol, solver ('plain'|'batched'|'fused'), precision (32|64),
rr_interval, use_graph, group_bytes_mb, verbose, autotune, and for tiled runs
stitch_integer (whole-cycle tile offsets) and merge_labels (connected
components labelled on the stitched scene)).  Unset keys follow the
reference defaults and its CUPHU_* environment overrides.
"""

from __future__ import annotations

import os

import numpy as np

from cuphu_amd._check import (
    check_bool_or_byte_dtype,
    check_complex_dtype,
    check_cost_mode,
    check_float_dtype,
    check_init_method,
    check_shape_match,
)
from cuphu_amd._ext import _cuphu_ext

__all__ = ["unwrap"]

# Default tile overlap for auto-tiled Laplace runs (_unwrap.py:28).
_DEFAULT_LAPLACE_TILE_OVERLAP = 64


def _auto_ntiles_by_target_size(nrow, ncol, target_tile_size, row_ovrlp, col_ovrlp):
    """(ntilerow, ntilecol) giving tiles of one shared edge length near
    target_tile_size (including overlap): the edge is taken from the larger
    dimension, rounded to a whole tile count, and reused on the other axis
    (_unwrap.py:31-61)."""
    ovrlp = max(row_ovrlp, col_ovrlp)
    step = max(1, target_tile_size - ovrlp)
    max_dim = max(nrow, ncol)
    n_max = max(1, round(max_dim / step))
    edge = max_dim / n_max
    return max(1, round(nrow / edge)), max(1, round(ncol / edge))


def _dilate_mask(valid: np.ndarray, pixels: int) -> np.ndarray:
    """Grow the True region by `pixels` rounds of 4-connectivity dilation
    (a Manhattan-distance diamond; _unwrap.py:64-84)."""
    m = valid
    for _ in range(pixels):
        grown = m.copy()
        grown[1:, :] |= m[:-1, :]
        grown[:-1, :] |= m[1:, :]
        grown[:, 1:] |= m[:, :-1]
        grown[:, :-1] |= m[:, 1:]
        m = grown
    return m


def unwrap(  # noqa: C901 - mirrors the reference's single entry point
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
    igram = np.asarray(igram)
    corr = np.asarray(corr)
    if igram.ndim != 2:
        raise ValueError(f"igram must be 2-D, got ndim={igram.ndim}")

    nrow, ncol = igram.shape
    check_shape_match((nrow, ncol), corr=corr)
    if mask is not None:
        mask = np.asarray(mask)
        check_shape_match((nrow, ncol), mask=mask)
    if mag is not None:
        mag = np.asarray(mag)
        check_shape_match((nrow, ncol), mag=mag)

    check_complex_dtype(igram=igram)
    check_float_dtype(corr=corr)
    if mask is not None:
        check_bool_or_byte_dtype(mask=mask)
    if mag is not None:
        check_float_dtype(mag=mag)

    check_cost_mode(cost)
    check_init_method(init)

    if nlooks < 1.0:
        raise ValueError(f"nlooks must be >= 1, got {nlooks}")

    if nproc < 1:
        nproc = os.cpu_count() or 1

    if tile_overlap is None:
        tile_overlap = _DEFAULT_LAPLACE_TILE_OVERLAP if init == "laplace" else 0
    if np.ndim(tile_overlap) == 0:
        tile_overlap = (int(tile_overlap), int(tile_overlap))
    row_ovrlp, col_ovrlp = tile_overlap

    if ntiles is None:
        if init == "laplace" or nproc > 1:
            ntilerow, ntilecol = _auto_ntiles_by_target_size(
                nrow, ncol, target_tile_size, row_ovrlp, col_ovrlp)
        else:
            ntilerow, ntilecol = 1, 1
    else:
        ntilerow, ntilecol = int(ntiles[0]), int(ntiles[1])

    # Non-finite interferogram elements and coherences count as 0
    # (_unwrap.py:515-523); the library applies that on the device after
    # the upload, so the arrays are passed as they are.
    igram_c64 = np.ascontiguousarray(igram, dtype=np.complex64)
    corr_f32 = np.ascontiguousarray(corr, dtype=np.float32)
    mask_u8 = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
    mag_f32 = np.ascontiguousarray(mag, dtype=np.float32) if mag is not None else None

    solve_mask_u8 = mask_u8
    if mask_u8 is not None and mask_pad_distance > 0:
        solve_mask_u8 = _dilate_mask(mask_u8 != 0, int(mask_pad_distance)).astype(np.uint8)

    # Neighbour feedback samples the seam pixels; with a padded solve mask it
    # runs after the solve on the ORIGINAL mask instead, so padded-through
    # pixels never count as boundary data.  single_tile_reoptimize supersedes
    # it (_unwrap.py:529-557).
    effective_neighbor_feedback = laplace_neighbor_feedback and not single_tile_reoptimize
    feedback_needs_orig_mask = (
        mask_u8 is not None and mask_pad_distance > 0 and effective_neighbor_feedback)
    internal_feedback = effective_neighbor_feedback and not feedback_needs_orig_mask

    kperpdpsi, kpardpsi = int(phase_grad_window[0]), int(phase_grad_window[1])

    unw_out, cc_out = _cuphu_ext.unwrap_arrays(
        igram_c64, corr_f32, float(nlooks),
        cost=cost,
        init=init,
        mask=solve_mask_u8,
        orig_mask=mask_u8,
        mag=mag_f32,
        kperpdpsi=kperpdpsi,
        kpardpsi=kpardpsi,
        min_conncomp_frac=float(min_conncomp_frac),
        ntilerow=ntilerow,
        ntilecol=ntilecol,
        tile_rowovrlp=row_ovrlp,
        tile_colovrlp=col_ovrlp,
        tilecostthresh=tile_cost_thresh,
        minregionsize=min_region_size,
        nproc=nproc,
        single_tile_reoptimize=bool(single_tile_reoptimize),
        laplace_neighbor_feedback=bool(internal_feedback),
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

    if feedback_needs_orig_mask:
        unw_out = _cuphu_ext._laplace_neighbor_feedback_test(
            np.ascontiguousarray(unw_out, dtype=np.float32), mask_u8, ntilerow, ntilecol,
            row_ovrlp, col_ovrlp, int(laplace_neighbor_feedback_feather))

    if mask_u8 is not None and mask_pad_distance > 0:
        # padded-through pixels are never reported valid (_unwrap.py:600-603)
        cc_out = np.where(mask_u8 != 0, cc_out, 0).astype(cc_out.dtype)

    if reference_pixel is not None:
        # integer-cycle shift so the reference pixel equals its own wrapped
        # phase in [0, 2pi) (_unwrap.py:605-617)
        rp, rc = int(reference_pixel[0]), int(reference_pixel[1])
        z = igram_c64[rp, rc]
        ref_phase = float(np.angle(z)) if np.isfinite(z) else 0.0   # a scrubbed element is 0 + 0i
        if ref_phase < 0.0:
            ref_phase += 2.0 * np.pi
        n = round((float(unw_out[rp, rc]) - ref_phase) / (2.0 * np.pi))
        if n != 0:
            unw_out = unw_out - np.float32(n * 2.0 * np.pi)

    if unw is not None:
        unw[...] = unw_out
        unw_out = unw
    if conncomp is not None:
        conncomp[...] = cc_out.astype(conncomp.dtype)
        cc_out = conncomp

    return unw_out, cc_out

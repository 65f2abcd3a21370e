"""Connected-component relabelling of an unwrapped phase
(reference: src/cuphu/_conncomp.py:12-66).

The reference's regrow_conncomp falls back to a CPU scipy implementation;
this one runs the GPU union-find of the unwrap path with the given
coherence floor.  Label order (increasing minimum raster index per
component) equals the reference's scipy order.
"""

from __future__ import annotations

import numpy as np

from cuphu_amd._ext import _cuphu_ext

__all__ = ["regrow_conncomp"]


def regrow_conncomp(
    unw: np.ndarray,
    corr: np.ndarray,
    *,
    mask: np.ndarray | None = None,
    cost_thresh: float = 0.1,
    min_conncomp_frac: float = 0.01,
    max_ncomps: int = 32,
    gpu_id: int = 0,
) -> np.ndarray:
    """Connected-component labels (uint32, 0 = unassigned) of *unw*: two
    4-neighbours join when both have corr >= cost_thresh (and mask != 0)
    and their phase difference is below pi."""
    unw = np.ascontiguousarray(unw, dtype=np.float32)
    corr = np.ascontiguousarray(corr, dtype=np.float32)
    mask_u8 = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
    return _cuphu_ext.regrow_conncomp_arrays(unw, corr, mask_u8, float(cost_thresh),
                                             float(min_conncomp_frac), int(max_ncomps), int(gpu_id))

"""cuphu_amd -- the Laplace PCG/IRLS phase-unwrapping path of cuPHU, built
for AMD Instinct MI355X (gfx950).

Quick start::

    import numpy as np
    import cuphu_amd as cuphu

    unw, conncomp = cuphu.unwrap(igram, corr, nlooks=10.0, init="laplace")

The public names mirror the reference package ``cuphu``
(src/cuphu/__init__.py): unwrap, regrow_conncomp, gpu_count, gpu_info,
gpu_name, get_effective_looks.  Additions: ``unwrap_batch`` (many
interferograms in one launch sequence, per-instance nlooks), ``WLSSolve``
(differentiable fixed-weight solve with an adjoint backward; needs torch)
and ``set_flow_solver`` (a host network-flow solver for init='mcf'/'mst',
around which the MCF path's GPU stages run).
"""

from cuphu_amd._unwrap import unwrap
from cuphu_amd._conncomp import regrow_conncomp
from cuphu_amd._ext import _cuphu_ext
from cuphu_amd._looks import get_effective_looks
from cuphu_amd.mcf import set_flow_solver

__version__ = "0.1.0"

# device queries, straight from the extension: gpu_count(), gpu_name(device_id=0),
# gpu_info(device_id=0) -> {name, total_memory, sm_count (compute units), compute_cap (gfx arch)}
gpu_count = _cuphu_ext.gpu_count
gpu_name = _cuphu_ext.gpu_name
gpu_info = _cuphu_ext.gpu_info

__all__ = [
    "__version__",
    "get_effective_looks",
    "gpu_count",
    "gpu_info",
    "gpu_name",
    "regrow_conncomp",
    "set_flow_solver",
    "unwrap",
    "unwrap_batch",
    "WLSSolve",
]


def __getattr__(name):
    # torch-backed additions are imported lazily
    if name == "unwrap_batch":
        from cuphu_amd.batch import unwrap_batch
        return unwrap_batch
    if name == "WLSSolve":
        from cuphu_amd.autograd import WLSSolve
        return WLSSolve
    raise AttributeError(name)

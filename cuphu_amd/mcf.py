"""Host network-flow solver hook for init='mcf' / 'mst'.

The reference runs SNAPHU's CPU network-flow code between its GPU stages
(src/cxx/cuphu/cuphu_solver.cpp:512-880: MCFInitFlows / MSTInitFlows and
TreeSolve).  That code is not part of this library.  ``set_flow_solver``
registers a Python callable in its place; ``unwrap(..., init='mcf')`` (one
tile) then runs the reference's GPU sequence around it: wrapped phase and
SNAPHU costs on the GPU, the callable for the integer flows, then on the
GPU the incremental costs at the solved flows, the flow integration and the
connected components gated by the incremental costs.

The callable receives ``(costs, phase, init)``:

* costs -- int16 array [narcs, 2] (smooth: offset, sigsq) or [narcs, 4]
  (defo/topo: offset, sigsq, dzmax, laycost), flat row-arc-then-column-arc
  order (narcs = (nrow-1)*ncol + nrow*(ncol-1));
* phase -- float32 wrapped phase [nrow, ncol] in [0, 2pi);
* init -- 'mcf' or 'mst';

and returns the integer flows, int16 [narcs] in the same arc order (row
arcs then column arcs, the layout ``cuphu_integrate_phase_gpu`` takes).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from cuphu_amd import _capi

__all__ = ["set_flow_solver", "incremental_costs"]

_INIT_NAMES = {v: k for k, v in _capi.INIT.items()}
_registered = None   # keeps the ctypes thunk alive while registered


def set_flow_solver(fn) -> None:
    """Register ``fn(costs, phase, init) -> flows`` (see module docstring) as
    the process-wide network-flow solver; ``None`` clears it."""
    global _registered
    if fn is None:
        _capi.check(_capi.lib.cuphu_set_flow_solver(None, None), "cuphu_set_flow_solver")
        _registered = None
        return

    def thunk(costs_p, cost_mode, phase_p, nrow, ncol, init_method, flows_p, user):
        try:
            narcs = (nrow - 1) * ncol + nrow * (ncol - 1)
            width = 2 if cost_mode == _capi.COST["smooth"] else 4
            costs = np.ctypeslib.as_array(C.cast(costs_p, C.POINTER(C.c_short)), shape=(narcs * width,))
            phase = np.ctypeslib.as_array(C.cast(phase_p, C.POINTER(C.c_float)), shape=(nrow * ncol,))
            flows = fn(costs.reshape(narcs, width).copy(), phase.reshape(nrow, ncol).copy(),
                       _INIT_NAMES.get(init_method, str(init_method)))
            flows = np.asarray(flows)
            if flows.shape != (narcs,):
                raise ValueError(f"flow solver returned shape {flows.shape}, expected ({narcs},)")
            out = np.ctypeslib.as_array(C.cast(flows_p, C.POINTER(C.c_short)), shape=(narcs,))
            out[:] = flows.astype(np.int16)
            return 0
        except Exception:   # noqa: BLE001 - reported to the library as a failed solve
            import traceback

            traceback.print_exc()
            return 1

    cb = _capi.FLOW_SOLVER_FN(thunk)
    _capi.check(_capi.lib.cuphu_set_flow_solver(C.cast(cb, C.c_void_p), None), "cuphu_set_flow_solver")
    _registered = cb


def incremental_costs(d_costs, d_flows, nshortcycle: int = 200, nflow: int = 1):
    """GPU incremental costs of smooth-cost arcs (cuphu_incrcost_gpu) on torch
    tensors: d_costs int16 [narcs, 2] (offset, sigsq), d_flows int16 [narcs].
    Returns (poscost, negcost, no_improvement)."""
    import torch

    n = int(d_flows.numel())
    pos = torch.empty(n, dtype=torch.int16, device=d_flows.device)
    neg = torch.empty(n, dtype=torch.int16, device=d_flows.device)
    flag = C.c_int()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check(_capi.lib.cuphu_incrcost_gpu(_capi.ptr(d_costs), _capi.ptr(d_flows), n, nshortcycle, nflow,
                                             _capi.ptr(pos), _capi.ptr(neg), C.byref(flag), stream),
                "cuphu_incrcost_gpu")
    return pos, neg, bool(flag.value)

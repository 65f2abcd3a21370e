"""One HIP runtime per process.

PyTorch-ROCm wheels bundle their own HIP and HSA runtimes
(torch/lib/libamdhip64.so, SONAME libamdhip64.so.7).  libcuphu_hip.so
needs libamdhip64.so.7 as well; loaded on its own it binds the system ROCm
runtime, and a later ``import torch`` would then start a second HSA
runtime that finds no GPU.  When torch is installed, its runtime is loaded
first (RTLD_GLOBAL) so that the library binds to that same instance and
torch tensors, streams and this library share one runtime.  Without torch
the system runtime is used.
"""

from __future__ import annotations

import ctypes
import importlib.util
import os

_done = False


def ensure_single_runtime() -> None:
    global _done
    if _done:
        return
    _done = True
    if os.environ.get("CUPHU_SYSTEM_HIP_RUNTIME"):
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)

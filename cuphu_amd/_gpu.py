"""GPU device query helpers (reference: src/cuphu/_gpu.py)."""

from cuphu_amd._ext import _cuphu_ext


def gpu_count() -> int:
    """Return the number of GPUs available on this machine."""
    return _cuphu_ext.gpu_count()


def gpu_name(device_id: int = 0) -> str:
    """Return the name string of the specified device."""
    return _cuphu_ext.gpu_name(device_id)


def gpu_info(device_id: int = 0) -> dict:
    """Return a dictionary of properties for the specified device
    (name, total_memory, sm_count = compute units, compute_cap = gfx arch)."""
    return _cuphu_ext.gpu_info(device_id)

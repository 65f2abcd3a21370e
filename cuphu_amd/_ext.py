"""Import of the compiled pybind11 extension (reference: src/cuphu/_ext.py).

There is no fallback: without the built HIP extension the package does not
import.
"""

from cuphu_amd._runtime import ensure_single_runtime

ensure_single_runtime()

try:
    import cuphu_amd._cuphu_ext as _cuphu_ext
except ImportError as e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "The cuphu_amd HIP extension module could not be imported. Build it for gfx950 with "
        "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root."
    ) from e

__all__ = ["_cuphu_ext"]

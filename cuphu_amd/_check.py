"""Argument checks of the Python front door.

The raster arguments are described by a table of rules (name, the dtype
class it must have, whether it may be absent) and checked in two sweeps:
structure first (2-D, the interferogram's shape), then element types.
The two option checks keep the names the reference's tests import
(``check_cost_mode`` and ``check_init_method``, src/cuphu/_check.py), and
every message keeps the reference's wording, which callers match on:
"must be 2-D", "shape ... does not match", "must be complex", "must be
floating-point", "cost must be one of", "init must be one of".
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

COST_MODES = ("smooth", "defo", "topo")
INIT_METHODS = ("mst", "mcf", "laplace")

# dtype class -> (test, words used in the message)
_DTYPE_CLASSES = {
    "complex": (lambda dt: np.issubdtype(dt, np.complexfloating), "complex"),
    "float": (lambda dt: np.issubdtype(dt, np.floating), "floating-point"),
    "integer": (lambda dt: np.issubdtype(dt, np.integer), "integer"),
    "mask": (lambda dt: dt in (np.dtype(np.bool_), np.dtype(np.uint8), np.dtype(np.int8)),
             "bool or 8-bit integer"),
}


class Raster(NamedTuple):
    """One 2-D argument: its name, value (None when absent) and dtype class."""
    name: str
    value: object
    kind: str


def _one_of(what: str, value: object, allowed: tuple[str, ...]) -> None:
    if value not in allowed:
        raise ValueError(f"{what} must be one of {set(allowed)}, got {value!r}")


def check_cost_mode(cost: str) -> None:
    _one_of("cost", cost, COST_MODES)


def check_init_method(init: str) -> None:
    _one_of("init", init, INIT_METHODS)


def check_dtype(name: str, arr: np.ndarray, kind: str) -> None:
    ok, words = _DTYPE_CLASSES[kind]
    if not ok(arr.dtype):
        raise TypeError(f"{name} must be {words}, got dtype={arr.dtype}")


def check_rasters(rules: list[Raster]) -> tuple[tuple[int, int], dict[str, np.ndarray]]:
    """Validate the rasters; the first rule's array fixes the shape.

    Returns (shape, {name: ndarray}) for the rasters that are present.
    """
    arrays = {r.name: np.asarray(r.value) for r in rules if r.value is not None}
    lead = rules[0].name
    if arrays[lead].ndim != 2:
        raise ValueError(f"{lead} must be 2-D, got ndim={arrays[lead].ndim}")
    shape = tuple(int(n) for n in arrays[lead].shape)
    for r in rules[1:]:
        if r.name in arrays and arrays[r.name].shape != shape:
            raise ValueError(f"{r.name} shape {arrays[r.name].shape} does not match expected {shape}")
    for r in rules:
        if r.name in arrays:
            check_dtype(r.name, arrays[r.name], r.kind)
    return shape, arrays

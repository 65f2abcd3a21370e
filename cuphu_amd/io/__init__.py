"""I/O dataset protocols (reference: src/cuphu/io/__init__.py)."""

from cuphu_amd.io._dataset import InputDataset, OutputDataset

__all__ = ["InputDataset", "OutputDataset"]

"""cudaem_amd: MI355X-native (gfx950) multislice engine.

The compute path is libcudaem_amd.so (hand-written HIP kernels behind the C
ABI of include/cudaem_amd.h); importing a GPU entry point without the built
library raises.  Python here is plumbing: torch tensors carry device memory.
"""
from ._capi import BORDER_MAX, BORDER_MIDDLE, BORDER_MIN, LENS_FIELDS, EmError  # noqa: F401
from .engine import Plan, fft_size_supported, ragged  # noqa: F401
from . import specimen  # noqa: F401

__all__ = ["Plan", "ragged", "fft_size_supported", "specimen", "EmError", "LENS_FIELDS",
           "BORDER_MIN", "BORDER_MAX", "BORDER_MIDDLE"]

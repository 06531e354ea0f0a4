"""proddl2_amd: MI355X-native FFT rigid-body correlation scan (PRODDL2's hot path).

The scan itself lives in libpdfft.so (HIP, gfx950) behind the C ABI in
include/pdfft.h; ``proddl2_amd.pdfft`` binds it, ``proddl2_amd.synth`` makes
seeded synthetic complexes, rotation sets and terms.
"""
from . import pdfft, synth  # noqa: F401
from .pdfft import PdError, Scan, find_best_size  # noqa: F401

__all__ = ["pdfft", "synth", "Scan", "PdError", "find_best_size"]

"""The reference's own gdiam (oracle/_ref/libgdiam_ref.so, built by `make -C
oracle ref` from the reference's vendored sources when they are present).
Test infrastructure only: a checker of pd_min_volume_box and pd_diameter_pair,
never part of the product.  `load()` returns None when the library was not
built (the reference sources are absent)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libgdiam_ref.so")
_lib = None


def load():
    global _lib
    if _lib is None and os.path.exists(_PATH):
        lib = C.CDLL(_PATH)
        d = C.POINTER(C.c_double)
        lib.gdref_mvbb.argtypes = [d, C.c_int, C.c_int, C.c_int, d, d, d]
        lib.gdref_mvbb.restype = None
        lib.gdref_diam.argtypes = [d, C.c_int, C.c_double, d, d]
        lib.gdref_diam.restype = None
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mvbb(pos, grid_n=5, sample_n=-1):
    """Geom::Bounding::Box(pos) as MolStruct::moveIntoMinBox builds it
    (bounding.hpp:135-183): (dirs[3, 3] rows, left corner, right corner), the
    corners in the input frame."""
    lib = load()
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    left, right, dirs = np.empty(3), np.empty(3), np.empty((3, 3))
    lib.gdref_mvbb(_dp(p), len(p), grid_n, sample_n, _dp(left), _dp(right), _dp(dirs))
    return dirs, left, right


def box_frame(pos, grid_n=5):
    """The reference's box transform and padded-box input
    (Box::getTransformationToBoundCoordinates, bounding.hpp:196-206):
    x -> dirs (x - centre); returns (dirs, centre, extents along dirs)."""
    dirs, left, right = mvbb(pos, grid_n)
    centre = left + (right - left) / 2
    ext = np.abs(dirs @ (right - left))
    return dirs, centre, ext


def diameter(pos, eps=0.0):
    lib = load()
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    a, b = np.empty(3), np.empty(3)
    lib.gdref_diam(_dp(p), len(p), float(eps), _dp(a), _dp(b))
    return a, b

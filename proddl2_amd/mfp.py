"""PDMF: the molforce-parameter container read by proddl-dock-fft.

It carries the datasets of the reference's molforce HDF5 file
(include/PRODDL/docking_io_hdf5.hpp:19-71; written by
python/lib/PRODDL/force_field_fft.py:24-101): pos (n,3), mass, eps, sigma,
alpha (n), mol_offsets (>=2, 2) and the mix attribute.  HDF5 is not available
in this image, so the same datasets go into a flat binary container:

    b"PDMF" | u32 version=1 | u32 count | count x entry
    entry: u32 name_len | name | u8 dtype ('d' f64 | 'i' i32) | u32 ndim |
           u64 shape[ndim] | data (C order, little endian)
"""
from __future__ import annotations

import struct

import numpy as np


def write(path: str, datasets: dict) -> None:
    with open(path, "wb") as f:
        f.write(b"PDMF")
        f.write(struct.pack("<II", 1, len(datasets)))
        for name, arr in datasets.items():
            a = np.asarray(arr)
            if a.dtype.kind == "f":
                a, code = np.ascontiguousarray(a, "<f8"), b"d"
            else:
                a, code = np.ascontiguousarray(a, "<i4"), b"i"
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)) + nb + code + struct.pack("<I", a.ndim))
            f.write(struct.pack(f"<{a.ndim}Q", *a.shape))
            f.write(a.tobytes())


def read(path: str) -> dict:
    raw = open(path, "rb").read()
    if raw[:4] != b"PDMF":
        raise ValueError("not a PDMF file")
    version, count = struct.unpack_from("<II", raw, 4)
    at = 12
    out = {}
    for _ in range(count):
        (nl,) = struct.unpack_from("<I", raw, at); at += 4
        name = raw[at:at + nl].decode(); at += nl
        code = raw[at:at + 1]; at += 1
        (nd,) = struct.unpack_from("<I", raw, at); at += 4
        shape = struct.unpack_from(f"<{nd}Q", raw, at); at += 8 * nd
        dt = "<f8" if code == b"d" else "<i4"
        n = int(np.prod(shape)) if nd else 1
        out[name] = np.frombuffer(raw, dt, n, at).reshape(shape); at += n * np.dtype(dt).itemsize
    return out


def write_molforce(path: str, rec_pos, lig_pos, mass, eps, sigma, alpha, mix: int = 1) -> None:
    """Receptor atoms first, then ligand atoms, as the reference's writer does."""
    nr, nl = len(rec_pos), len(lig_pos)
    write(path, {
        "pos": np.concatenate([rec_pos, lig_pos]).astype(float),
        "mass": np.asarray(mass, float), "eps": np.asarray(eps, float),
        "sigma": np.asarray(sigma, float), "alpha": np.asarray(alpha, float),
        "mol_offsets": np.array([[0, nr], [nr, nr + nl]], np.int32),
        "mix": np.array([mix], np.int32),
    })

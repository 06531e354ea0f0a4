"""CPU oracle of the Foreman gather (test infrastructure only).

Restates, in FP64 numpy, what `proddl-dock-fft --task gather` computes from
the rotation-scan files of the workers:

  - molecule frames (MolStruct::randomizeLigandOrientation / moveIntoMinBox,
    include/PRODDL/docking_molstruct.hpp:199-266) for a given receptor box:
    both molecules moved by the box transform (axes in rows, centre to the
    origin), the ligand first spun about its centre of mass by startingSpin and
    finally moved so that the centre of its diameter is at the origin;
  - Foreman::run (include/PRODDL/docking_app.hpp:174-220): every translation of
    every record becomes the rigid transformation Translation(tran) * Rotation(rot);
    the maxNRigidMatches lowest values below 1e16 are kept, ascending;
  - Foreman::getResults (:245-262) through ToOriginalFrameTransformer
    (docking_molstruct.hpp:31-51): tR^-1 * t * tL;
  - IORigid::writeCoords, binary format (include/PRODDL/IO/rigid.hpp:29-150).

Among equal values the reference's order comes from the libstdc++ heap; the
tests use distinct values, for which the kept set and its order are fixed.
Only tests/ may import this module.
"""
from __future__ import annotations

import struct

import numpy as np

from . import pdoracle

IORIGID_SIGNATURE = 2005052501


def rt(m=None, v=None):
    return (np.eye(3) if m is None else np.asarray(m, np.float64).reshape(3, 3),
            np.zeros(3) if v is None else np.asarray(v, np.float64))


def compose(a, b):
    return a[0] @ b[0], a[0] @ b[1] + a[1]


def inverse(a):
    mt = a[0].T
    return mt, mt @ (-a[1])


def _apply(t, x):
    """x -> m x + v row by row, summed in the order the CLI sums (IEEE per op)."""
    m, v = t
    x = np.asarray(x, np.float64)
    return np.stack([((m[i, 0] * x[:, 0] + m[i, 1] * x[:, 1]) + m[i, 2] * x[:, 2]) + v[i] for i in range(3)], axis=1)


def docking_frame(rec_pos, lig_pos, lig_mass, spin, box, h=None, cut=None):
    """The docking frame (MolStruct::moveIntoReferencePositions,
    docking_molstruct.hpp:199-281) for a given receptor box (axes in rows, lo,
    hi along them): the ligand spun about its centre of mass, both molecules
    moved by x -> A x - (lo + hi) / 2, the ligand then centred on its diameter.
    Returns (receptor, ligand, tR, tL, ligand size) and, given the grid step and
    cutoff, the grid edges of the padded box (:232-243)."""
    axes, blo, bhi = (np.asarray(a, np.float64) for a in box)
    rs = pdoracle.euler_to_matrix([spin, spin * 3.7877, spin * -4.513432])
    cm = (lig_mass[:, None] * lig_pos).sum(0) / lig_mass.sum()
    spun = _apply((rs, np.zeros(3)), lig_pos - cm) + cm
    tb = (axes.reshape(3, 3), -0.5 * (blo + bhi))
    rec_d = _apply(tb, rec_pos)
    lig_b = _apply(tb, spun)
    d2 = ((lig_b[:, None, :] - lig_b[None, :, :]) ** 2).sum(-1)
    i, j = np.unravel_index(np.argmax(np.triu(d2, 1)), d2.shape)
    mid = 0.5 * (lig_b[i] + lig_b[j])
    lig_d = lig_b - mid
    tR = tb
    spin_cm = compose(compose(rt(v=cm), rt(m=rs)), rt(v=-cm))
    tL = compose(rt(v=-mid), compose(tb, spin_cm))
    size = float(np.sqrt(d2[i, j]))
    out = (rec_d, lig_d, tR, tL, size)
    if h is None:
        return out
    pad = 2 * h + size / 2 + cut
    half = 0.5 * (bhi - blo)
    n = tuple(pdoracle.grid_size_from_box(-half[k] - pad, half[k] + pad, h) for k in range(3))
    return out + (n,)


def frames(rec_pos, lig_pos, lig_mass, spin, box):
    """(tR, tL): original -> docking frame of receptor and ligand."""
    _, _, tR, tL, _ = docking_frame(rec_pos, lig_pos, lig_mass, spin, box)
    return tR, tL


def gather(scan_files, rec_pos, lig_pos, lig_mass, spin, max_matches, box):
    """List of (value, rotation matrix, translation) in the original frame, ascending."""
    items = []
    for path in scan_files:
        rots, counts, tran, vals = pdoracle.read_scan_records(path)
        if len(rots) == 0:
            break  # the collector stops at the first file without a record
        for r in range(len(rots)):
            rot = rots[r].reshape(3, 3)
            for k in range(counts[r]):
                if vals[r, k] < 1e16:
                    items.append((float(vals[r, k]), compose(rt(v=tran[r, k]), rt(m=rot))))
    items.sort(key=lambda x: x[0])
    items = items[:max_matches]
    tR, tL = frames(rec_pos, lig_pos, lig_mass, spin, box)
    tRi = inverse(tR)
    return [(v, *compose(compose(tRi, t), tL)) for v, t in items]


def read_iorigid(path):
    raw = open(path, "rb").read()
    sig, n = struct.unpack_from("<ii", raw, 0)
    if sig != IORIGID_SIGNATURE:
        raise ValueError("not an IORigid file")
    rec = np.frombuffer(raw, "<f4", count=7 * n, offset=8).reshape(n, 7)
    return rec[:, 0], rec[:, 1:4], rec[:, 4:7]

"""CPU oracle of the Foreman gather (test infrastructure only).

Restates, in FP64 numpy, what `proddl-dock-fft --task gather` computes from
the rotation-scan files of the workers:

  - molecule frames (MolStruct::randomizeLigandOrientation / moveIntoMinBox,
    include/PRODDL/docking_molstruct.hpp:199-266): the receptor moved by its box
    transform (here the axis-aligned box centred at the origin), the ligand spun
    about its centre of mass by startingSpin and moved so that the centre of its
    diameter is at the origin;
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


def frames(rec_pos, lig_pos, lig_mass, spin):
    """(tR, tL): original -> docking frame of receptor and ligand."""
    rs = pdoracle.euler_to_matrix([spin, spin * 3.7877, spin * -4.513432])
    cm = (lig_mass[:, None] * lig_pos).sum(0) / lig_mass.sum()
    spun = (lig_pos - cm) @ rs.T + cm
    lo, hi = rec_pos.min(0), rec_pos.max(0)
    c = 0.5 * (lo + hi)
    d2 = ((spun[:, None, :] - spun[None, :, :]) ** 2).sum(-1)
    i, j = np.unravel_index(np.argmax(np.triu(d2, 1)), d2.shape)
    mid = 0.5 * (spun[i] + spun[j])
    tR = rt(v=-c)
    tL = compose(rt(v=-mid), compose(compose(rt(v=cm), rt(m=rs)), rt(v=-cm)))
    return tR, tL


def gather(scan_files, rec_pos, lig_pos, lig_mass, spin, max_matches):
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
    tR, tL = frames(rec_pos, lig_pos, lig_mass, spin)
    tRi = inverse(tR)
    return [(v, *compose(compose(tRi, t), tL)) for v, t in items]


def read_iorigid(path):
    raw = open(path, "rb").read()
    sig, n = struct.unpack_from("<ii", raw, 0)
    if sig != IORIGID_SIGNATURE:
        raise ValueError("not an IORigid file")
    rec = np.frombuffer(raw, "<f4", count=7 * n, offset=8).reshape(n, 7)
    return rec[:, 0], rec[:, 1:4], rec[:, 4:7]

"""Writes tests/golden/frame_2ptc.npz: the atom coordinates of the reference's
own test complex (test_data/pdb/2ptc_E.pdb, the receptor, and 2ptc_I.pdb, the
ligand) and what the reference's vendored gdiam computes for them, the
receptor's minimum-volume box (Geom::Bounding::Box as MolStruct::moveIntoMinBox
builds it, grid 5 over all points) and both molecules' diameter pairs
(Geom::Bounding::Diameter, eps = 0).

Run in a checkout that has the reference sources, after `make -C oracle ref`:

    python tests/golden/make_frame_golden.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gdiam_ref  # noqa: E402


def pdb_coordinates(path):
    """x, y, z of every ATOM / HETATM record (PDB columns 31-54)."""
    out = []
    with open(path) as f:
        for line in f:
            if line.startswith(("ATOM  ", "HETATM")):
                out.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    return np.array(out, np.float64)


def main(ref):
    if gdiam_ref.load() is None:
        raise SystemExit("oracle/_ref/libgdiam_ref.so is missing: run make -C oracle ref")
    rec = pdb_coordinates(os.path.join(ref, "test_data", "pdb", "2ptc_E.pdb"))
    lig = pdb_coordinates(os.path.join(ref, "test_data", "pdb", "2ptc_I.pdb"))
    dirs, left, right = gdiam_ref.mvbb(rec)
    rd = gdiam_ref.diameter(rec)
    ld = gdiam_ref.diameter(lig)
    np.savez_compressed(os.path.join(HERE, "frame_2ptc.npz"), receptor=rec, ligand=lig, box_dirs=dirs,
                        box_left=left, box_right=right, rec_diam=np.stack(rd), lig_diam=np.stack(ld))
    print(f"receptor {len(rec)} atoms, ligand {len(lig)} atoms; box extents "
          f"{np.abs(dirs @ (right - left))}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

"""`proddl-dock-fft --task gather` (the Foreman, docking_app.hpp:174-281) on the CPU:
rotation-scan records in, IORigid matches out, checked against oracle/foreman.py
and against the closed form of the frame transform."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import foreman, pdoracle
from proddl2_amd import pdfft, mfp, synth
from proddl2_amd.build import CLI


def _molforce(path, rec, lig, rng):
    pos = np.concatenate([rec, lig])
    n = len(pos)
    mfp.write(path, {
        "pos": pos, "mass": rng.uniform(10, 16, n), "eps": rng.uniform(0.1, 0.3, n),
        "sigma": rng.uniform(3.0, 4.0, n), "alpha": rng.uniform(0.0, 0.5, n),
        "mol_offsets": np.array([[0, len(rec)], [len(rec), n]], np.int32), "mix": np.array([1], np.int32),
    })
    return pos


def _scan_files(tmp_path, rng, n_files=3, n_rec=4, K=6):
    paths = []
    for f in range(n_files):
        p = str(tmp_path / f"scan_{f}.bin")
        for r in range(n_rec):
            rot = synth.random_rotation(rng)
            tran = rng.uniform(-20, 20, (K, 3))
            vals = rng.normal(-100, 30, K)
            if r == 1:
                vals[0] = 2e16  # above the value limit: never kept
            pdoracle.write_scan_record(p, rot, tran, vals, append=True)
        paths.append(p)
    return paths


def _run(tmp_path, rec, lig, spin, max_matches, seed=0):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    rng = np.random.default_rng(seed)
    mf = str(tmp_path / "mf.pdmf")
    _molforce(mf, rec, lig, rng)
    masses = mfp.read(mf)["mass"][len(rec):]
    scans = _scan_files(tmp_path, rng)
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(scans) + "\n")
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"startingSpin": spin, "maxNRigidMatches": max_matches, "gridStep": 1.5,
                                "cutOffFft": 8.0}))
    out = str(tmp_path / "res.bin")
    subprocess.run([CLI, "--task", "gather", "--options", str(opts), "--molforce-params", mf,
                    "--fft-rot-scan-list", str(lst), "--fft-res", out], check=True, capture_output=True)
    return foreman.read_iorigid(out), scans, masses


def _molecules(seed=1, cuboid=False):
    """Receptor and ligand points; with cuboid, the receptor holds the corners of
    [-15, 25] x [-10, 20] x [-5, 15], which are then its minimum box with the
    coordinate axes as its axes."""
    rng = np.random.default_rng(seed)
    if not cuboid:
        return rng.uniform(-15, 25, (60, 3)), rng.uniform(-5, 9, (25, 3))
    lo, ext = np.array([-15.0, -10.0, -5.0]), np.array([40.0, 30.0, 20.0])
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)])
    rec = lo + np.concatenate([corners, rng.uniform(0, 1, (52, 3))]) * ext
    return rec, rng.uniform(-5, 9, (25, 3))


def test_gather_matches_oracle(tmp_path):
    rec, lig = _molecules()
    (val, ang, xyz), scans, masses = _run(tmp_path, rec, lig, spin=23.3734432, max_matches=40)
    want = foreman.gather(scans, rec, lig, masses, 23.3734432, 40, pdfft.min_volume_box(rec))
    assert len(val) == len(want) == 40
    for i, (v, m, t) in enumerate(want):
        assert val[i] == np.float32(v)
        # compare rotations as matrices (angle triples have branch cuts)
        got_m = pdoracle.euler_to_matrix(ang[i].astype(np.float64))
        assert np.abs(got_m - m).max() < 2e-6, i
        assert np.abs(xyz[i] - t).max() < 2e-5 * max(1.0, np.abs(t).max()), i
    assert np.all(np.diff(val) >= 0)


def test_gather_frame_transform_closed_form(tmp_path):
    """startingSpin = 0 and a receptor whose minimum box is axis-parallel:
    tL = Translation(-mid), tR = Translation(-c), so a match t = (R, s) maps to
    rotation R and translation c + s - R mid."""
    rec, lig = _molecules(2, cuboid=True)
    axes, _, _ = pdfft.min_volume_box(rec)
    assert np.abs(axes - np.eye(3)).max() < 1e-12
    (val, ang, xyz), scans, _ = _run(tmp_path, rec, lig, spin=0.0, max_matches=1000, seed=3)
    c = np.array([5.0, 5.0, 5.0])
    d2 = ((lig[:, None] - lig[None]) ** 2).sum(-1)
    i, j = np.unravel_index(np.argmax(d2), d2.shape)
    mid = 0.5 * (lig[i] + lig[j])
    expect = []
    for p in scans:
        rots, counts, tran, vals = pdoracle.read_scan_records(p)
        for r in range(len(rots)):
            R = rots[r].reshape(3, 3)
            for k in range(counts[r]):
                if vals[r, k] < 1e16:
                    expect.append((vals[r, k], R, c + tran[r, k] - R @ mid))
    expect.sort(key=lambda x: x[0])
    assert len(val) == len(expect) == 3 * 4 * 6 - 3
    for k, (v, R, t) in enumerate(expect):
        assert val[k] == np.float32(v)
        assert np.abs(pdoracle.euler_to_matrix(ang[k].astype(np.float64)) - R).max() < 2e-6
        assert np.abs(xyz[k] - t).max() < 1e-4


def test_gather_keeps_the_lowest_bounded(tmp_path):
    rec, lig = _molecules(4)
    (val, _, _), scans, masses = _run(tmp_path, rec, lig, spin=1.5, max_matches=7, seed=5)
    allv = np.concatenate([pdoracle.read_scan_records(p)[3][:, :6].reshape(-1) for p in scans])
    assert np.array_equal(val, np.sort(allv[allv < 1e16])[:7].astype(np.float32))


@pytest.mark.parametrize("option,value,what", [
    ("potentialName", "ljFoo", "potentialName"),
])
def test_rot_scan_refuses_what_it_cannot_do(tmp_path, option, value, what):
    """An option value the reference's App::init rejects (docking_app.hpp:56-67)
    ends the task with not_supported_error before any device work."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    rng = np.random.default_rng(4)
    cx = synth.make_complex(60, 20, 4)
    mf = str(tmp_path / "mf.pdmf")
    _molforce(mf, cx.receptor.pos, cx.ligand.pos, rng)
    ang = tmp_path / "ang.txt"
    ang.write_text("1\n0 0 0\n")
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"gridStep": 1.5, "cutOffFft": 8.0, "anglesFile": str(ang), option: value}))
    r = subprocess.run([CLI, "--task", "rot-scan", "--options", str(opts), "--molforce-params", mf,
                        "--fft-rot-grid-start", "0", "--fft-rot-grid-end", "1", "--fft-rot-scan-res",
                        str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "not_supported_error" in r.stderr and what in r.stderr, r.stderr
    assert not os.path.exists(tmp_path / "o.bin")

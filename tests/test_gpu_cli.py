"""The worker flow of the reference's Makeflow rules (python/lib/PRODDL/dock.py:86-113)
through this build's proddl-dock-fft: `rot-scan` tasks over rotation ranges, then
`gather` over their outputs.  The rot-scan records are compared with the FP64
oracle's scan of the same rotations in the same reference frame."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import foreman, pdoracle
from proddl2_amd import mfp, pdfft, synth
from proddl2_amd.build import CLI

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SPIN = 23.3734432  # startingSpin default (docking_app.hpp via docking_molstruct.hpp:248-266)


def _inputs(tmp_path, n_rec=400, n_lig=90, seed=12, n_ang=40):
    rng = np.random.default_rng(11)
    cx = synth.make_complex(n_rec, n_lig, seed)
    rec, lig = cx.receptor.pos + 7.0, cx.ligand.pos - 3.0  # arbitrary original frames
    pos = np.concatenate([rec, lig])
    n = len(pos)
    p = {"mass": rng.uniform(10, 16, n), "eps": rng.uniform(0.1, 0.3, n), "sigma": rng.uniform(3.0, 4.0, n),
         "alpha": rng.uniform(0.2, 0.5, n)}
    mf = str(tmp_path / "mf.pdmf")
    mfp.write(mf, {"pos": pos, **p, "mol_offsets": np.array([[0, len(rec)], [len(rec), n]], np.int32),
                   "mix": np.array([1], np.int32)})
    ang = tmp_path / "angles.txt"
    a = rng.uniform(-180, 180, (n_ang, 3))
    ang.write_text(f"{n_ang}\n" + "\n".join(f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in a) + "\n")
    return rec, lig, p, mf, str(ang), a


def _docking_frame(rec, lig, lig_mass, h, cut):
    """The rot-scan frame (docking_molstruct.hpp:199-266 as the CLI states it):
    both molecules moved into the receptor's minimum-volume box (the library's
    pd_min_volume_box, checked against the reference's gdiam in
    tests/test_frame.py), the ligand spun about its centre of mass and its
    diameter centred at the origin; grid edges from the padded box."""
    box = pdfft.min_volume_box(rec)
    rec_d, lig_d, tR, tL, size, n = foreman.docking_frame(rec, lig, lig_mass, SPIN, box, h, cut)
    return rec_d, lig_d, n, tR, tL


def _rot_scan(tmp_path, opts, mf, lo, hi):
    out = str(tmp_path / f"scan_{lo}.bin")
    print(f"rot-scan [{lo},{hi})", flush=True)
    r = subprocess.run([CLI, "--task", "rot-scan", "--options", str(opts), "--molforce-params", mf,
                        "--fft-rot-grid-start", str(lo), "--fft-rot-grid-end", str(hi), "--fft-rot-scan-res", out],
                       capture_output=True, timeout=120, text=True)
    assert r.returncode == 0, r.stderr
    print("  " + r.stdout.strip(), flush=True)
    return out


def _check_against_oracle(oracle_terms, rec, lig, p, n_rec, h, cut, rots, tran, vals, K):
    lig_mass = p["mass"][n_rec:]
    rec_d, lig_d, n, _, _ = _docking_frame(rec, lig, lig_mass, h, cut)
    rt, w = oracle_terms(p, n_rec)
    R = np.stack([pdoracle.project_receptor(n, h, rec_d, k, par, c, md) for (k, c, md, par) in rt])
    S = pdoracle.receptor_spectra(R)
    from tests.parity_helpers import tolerance
    tol = tolerance(R, w, n)
    nout_o, off_o, s_o, val_o = pdoracle.scan(n, h, S, lig_d, w, rots, K=K + 1, max_val=1e16, n_threads=4)
    exact = 0
    for r in range(len(rots)):
        v = val_o[r]
        for i in range(K):
            lo_gap = v[i] - v[i - 1] if i else np.inf
            if lo_gap > 2 * tol and v[i + 1] - v[i] > 2 * tol:
                assert np.array_equal(tran[r, i], h * s_o[r, i]), (r, i, tran[r, i], s_o[r, i])
                assert abs(vals[r, i] - v[i]) <= 1e-9 * max(1.0, abs(v[i])), (r, i, vals[r, i], v[i])
                exact += 1
            assert abs(vals[r, i] - v[i]) <= 3 * tol
    print(f"  grid {n}: {exact} of {len(rots) * K} ranks exact against the oracle (tol={tol:.3e})", flush=True)
    assert exact >= len(rots) * K // 2
    return n


def _ljcomp_terms(p, nr):
    rep, att, wr, wa = pdoracle.ljcomp_params(p["eps"][:nr], p["sigma"][:nr], p["alpha"][:nr], p["eps"][nr:],
                                             p["sigma"][nr:])
    return [(pdfft.FIELD_LJ_REP, 8.0, -1.0, rep), (pdfft.FIELD_LJ_ATT, 8.0, -1.0, att)], np.stack([wr, wa])


def _ljavg_terms(p, nr):
    par, w = pdoracle.ljavg_params(p["eps"][:nr], p["sigma"][:nr], p["alpha"][:nr], p["eps"][nr:], p["sigma"][nr:],
                                   1, 8.0)
    return [(pdfft.FIELD_LJ_FULL, 8.0, -1.0, par)], w.reshape(1, -1)


def test_rot_scan_ranges_then_gather(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    rec, lig, p, mf, ang_path, a = _inputs(tmp_path)
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"gridStep": 1.5, "cutOffFft": 8.0, "anglesFile": ang_path, "maxNTrans": 5,
                                "maxValCorr": 1e16, "maxNRigidMatches": 30}))
    outs = []
    for lo, hi in ((0, 25), (25, 40)):
        out = _rot_scan(tmp_path, opts, mf, lo, hi)
        rots, counts, tran, vals = pdoracle.read_scan_records(out)
        assert len(rots) == hi - lo and (counts == 5).all()
        assert np.all(np.diff(vals[:, :5], axis=1) >= 0)
        # each record's rotation is the angles file's rotation of that index
        for r in range(len(rots)):
            want = pdoracle.euler_to_matrix(np.radians(a[lo + r]))
            assert np.abs(rots[r].reshape(3, 3) - want).max() < 1e-12
        _check_against_oracle(_ljcomp_terms, rec, lig, p, len(rec), 1.5, 8.0, rots, tran[:, :5], vals[:, :5], 5)
        outs.append(out)
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(outs) + "\n")
    res = str(tmp_path / "res.bin")
    subprocess.run([CLI, "--task", "gather", "--options", str(opts), "--molforce-params", mf,
                    "--fft-rot-scan-list", str(lst), "--fft-res", res], check=True, capture_output=True, timeout=120)
    val, ang_o, xyz = foreman.read_iorigid(res)
    want = foreman.gather(outs, rec, lig, p["mass"][len(rec):], SPIN, 30, pdfft.min_volume_box(rec))
    assert len(val) == 30
    for i, (v, m, t) in enumerate(want):
        assert val[i] == np.float32(v)
        assert np.abs(pdoracle.euler_to_matrix(ang_o[i].astype(np.float64)) - m).max() < 2e-6
        assert np.abs(xyz[i] - t).max() < 1e-4


def test_rot_scan_ljavg_matches_oracle(tmp_path):
    """potentialName ljAvg (App::init, docking_app.hpp:60-64): one LJ_FULL term."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rec, lig, p, mf, ang_path, a = _inputs(tmp_path, seed=13, n_ang=12)
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"gridStep": 1.5, "cutOffFft": 8.0, "anglesFile": ang_path, "maxNTrans": 6,
                                "maxValCorr": 1e16, "potentialName": "ljAvg"}))
    out = _rot_scan(tmp_path, opts, mf, 0, 12)
    rots, counts, tran, vals = pdoracle.read_scan_records(out)
    assert len(rots) == 12 and (counts == 6).all()
    _check_against_oracle(_ljavg_terms, rec, lig, p, len(rec), 1.5, 8.0, rots, tran[:, :6], vals[:, :6], 6)


def test_rot_scan_postfilters_match_library(tmp_path):
    """doClusterTranslations and nMultimer (Worker::init, docking_app.hpp:363-398)
    through the CLI: the records equal the library's post-filtered scan of the
    same rotations in the same frame (tR^-1 and tL of the frame, half the
    ligand's diameter, clusterRadiusTrans, maxRmsdSymm)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rec, lig, p, mf, ang_path, a = _inputs(tmp_path, seed=14, n_ang=10)
    h, cut, K, k_inp = 1.5, 8.0, 6, 400
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"gridStep": h, "cutOffFft": cut, "anglesFile": ang_path, "maxNTrans": K,
                                "maxNTransInp": k_inp, "maxValCorr": 1e16, "doClusterTranslations": True,
                                "clusterRadiusTrans": 4.0, "nMultimer": 2, "maxRmsdSymm": 30.0,
                                "rescore64": False}))
    out = _rot_scan(tmp_path, opts, mf, 0, 10)
    rots, counts, tran, vals = pdoracle.read_scan_records(out)
    nr = len(rec)
    lig_mass = p["mass"][nr:]
    rec_d, lig_d, n, tR, tL = _docking_frame(rec, lig, lig_mass, h, cut)
    ai = foreman.inverse(tR)
    rep, att, wr, wa = pdfft.ljcomp_params(p["eps"][:nr], p["sigma"][:nr], p["alpha"][:nr], p["eps"][nr:],
                                           p["sigma"][nr:])
    d2 = ((lig_d[:, None, :] - lig_d[None, :, :]) ** 2).sum(-1)
    print(f"  grid {n}", flush=True)
    with pdfft.Scan(n, h, 2, max_n_trans=K, max_n_trans_inp=k_inp, max_val_corr=1e16) as sc:
        sc.project_receptor(rec_d, [(pdfft.FIELD_LJ_REP, cut, -1.0, rep), (pdfft.FIELD_LJ_ATT, cut, -1.0, att)])
        sc.set_ligand(lig_d, np.stack([wr, wa]))
        sc.set_postfilter(n_multimer=2, max_rmsd_symm=30.0, ligand_radius=np.sqrt(d2.max()) / 2,
                          frame_rec_inv=np.concatenate([ai[0].reshape(9), ai[1]]),
                          frame_lig=np.concatenate([tL[0].reshape(9), tL[1]]), cluster=True, cluster_radius=4.0)
        peaks, nout = sc.run(rots.reshape(-1, 9))
    same = 0
    for r in range(len(rots)):
        if counts[r] != nout[r]:
            continue
        k = int(nout[r])
        if np.array_equal(tran[r, :k], h * peaks["s"][r, :k]) and np.array_equal(vals[r, :k], peaks["value"][r, :k]):
            same += 1
    print(f"  {same} of {len(rots)} rotations identical; outputs per rotation {list(counts)}", flush=True)
    assert same >= len(rots) - 1
    assert 0 < counts.sum() < K * len(rots) + 1

"""The worker flow of the reference's Makeflow rules (python/lib/PRODDL/dock.py:86-113)
through this build's proddl-dock-fft: two `rot-scan` tasks over rotation ranges,
then `gather` over their outputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import foreman, pdoracle
from proddl2_amd import mfp, synth
from proddl2_amd.build import CLI

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_rot_scan_ranges_then_gather(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    rng = np.random.default_rng(11)
    cx = synth.make_complex(400, 90, 12)
    rec, lig = cx.receptor.pos + 7.0, cx.ligand.pos - 3.0  # arbitrary original frames
    pos = np.concatenate([rec, lig])
    n = len(pos)
    masses = rng.uniform(10, 16, n)
    mf = str(tmp_path / "mf.pdmf")
    mfp.write(mf, {"pos": pos, "mass": masses, "eps": rng.uniform(0.1, 0.3, n),
                   "sigma": rng.uniform(3.0, 4.0, n), "alpha": rng.uniform(0.2, 0.5, n),
                   "mol_offsets": np.array([[0, len(rec)], [len(rec), n]], np.int32),
                   "mix": np.array([1], np.int32)})
    ang = tmp_path / "angles.txt"
    a = rng.uniform(-180, 180, (40, 3))
    ang.write_text("40\n" + "\n".join(f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in a) + "\n")
    opts = tmp_path / "opt.json"
    opts.write_text(json.dumps({"gridStep": 1.5, "cutOffFft": 8.0, "anglesFile": str(ang), "maxNTrans": 5,
                                "maxValCorr": 1e16, "maxNRigidMatches": 30}))
    outs = []
    for lo, hi in ((0, 25), (25, 40)):
        out = str(tmp_path / f"scan_{lo}.bin")
        print(f"rot-scan [{lo},{hi})", flush=True)
        subprocess.run([CLI, "--task", "rot-scan", "--options", str(opts), "--molforce-params", mf,
                        "--fft-rot-grid-start", str(lo), "--fft-rot-grid-end", str(hi), "--fft-rot-scan-res", out],
                       check=True, capture_output=True, timeout=120)
        rots, counts, tran, vals = pdoracle.read_scan_records(out)
        assert len(rots) == hi - lo and (counts == 5).all()
        assert np.all(np.diff(vals[:, :5], axis=1) >= 0)
        # each record's rotation is the angles file's rotation of that index
        for r in range(len(rots)):
            want = pdoracle.euler_to_matrix(np.radians(a[lo + r]))
            assert np.abs(rots[r].reshape(3, 3) - want).max() < 1e-12
        outs.append(out)
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(outs) + "\n")
    res = str(tmp_path / "res.bin")
    subprocess.run([CLI, "--task", "gather", "--options", str(opts), "--molforce-params", mf,
                    "--fft-rot-scan-list", str(lst), "--fft-res", res], check=True, capture_output=True, timeout=120)
    val, ang_o, xyz = foreman.read_iorigid(res)
    want = foreman.gather(outs, rec, lig, masses[len(rec):], 23.3734432, 30)
    assert len(val) == 30
    for i, (v, m, t) in enumerate(want):
        assert val[i] == np.float32(v)
        assert np.abs(pdoracle.euler_to_matrix(ang_o[i].astype(np.float64)) - m).max() < 2e-6
        assert np.abs(xyz[i] - t).max() < 1e-4

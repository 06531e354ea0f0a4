"""Launch choices for the autotuner (pd_scan_set_tuning / pd_scan_get_tuning).

Column groups and the batch lanes must leave the results bit-identical.  The
column-pass variant (a larger one may run a first DFT stage that skips zero
planes) and the ligand pass (whole-plane or row and column kernels) can run
other arithmetic, so their correlation grids are held to the FP32 bar against
the oracle and their peaks to the oracle outside the band, like any other
path."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from proddl2_amd import pdfft, synth  # noqa: E402

from tests.parity_helpers import build_case, check_grids, compare_rotation_peaks, make_scan, tolerance  # noqa: E402

INVALID, UNSUPPORTED, STATE = 1, 5, 6  # pd_status


def _run(sc, rots):
    peaks, nout = sc.run(rots)
    return peaks.view(np.uint8).copy(), nout.copy()


@pytest.mark.parametrize("n,h,n_rec,n_lig", [((96, 96, 96), 1.2, 2500, 700), ((144, 144, 144), 1.0, 3000, 800)])
def test_bitwise_choices_keep_results(oracle, n, h, n_rec, n_lig):
    print(f"case tuning bit-identity n={n}", flush=True)
    cx, rt, w, _ = build_case(oracle, n, h, n_rec, n_lig, 4, seed=41)
    rots = synth.rotation_set(150, 42)
    with make_scan(n, h, 4, 8, cx, rt, w, batch=64) as sc:
        ref = _run(sc, rots)
        base = sc.tuning()
        print(f"  rules: {base}", flush=True)
        # lanes: two by default up to N0 = 128, three above
        assert base["col_groups"] >= 1 and base["lanes"] == (3 if n[0] > 128 else 2) and base["col_mz_extra"] == 0
        for kw in ({"col_groups": 1}, {"col_groups": 2}, {"col_groups": 16}, {"col_groups": 64},
                   {"lanes": 1}, {"lanes": 2}, {"lanes": 3}, {"col_groups": 8, "lanes": 2}):
            sc.set_tuning(**kw)
            eff = sc.tuning()
            if "lanes" in kw:
                assert eff["lanes"] == kw["lanes"], (kw, eff)
            got = _run(sc, rots)
            assert np.array_equal(got[1], ref[1]) and np.array_equal(got[0], ref[0]), kw
            print(f"  {kw}: effective {eff}, identical", flush=True)
        sc.set_tuning()
        assert sc.tuning() == base
        assert all(np.array_equal(a, b) for a, b in zip(_run(sc, rots), ref))


@pytest.mark.parametrize("n,h,n_rec,n_lig", [((96, 96, 96), 1.2, 2500, 700), ((144, 144, 144), 1.0, 3000, 800)])
def test_arithmetic_choices_match_oracle(oracle, n, h, n_rec, n_lig):
    print(f"case ligand pass and column variants n={n}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, 4, seed=43)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(6, 44)
    K = 8
    tol = tolerance(R, w, n)
    nout_o, off_o, s_o, val_o = oracle.scan(n, h, S, cx.ligand.pos, w, rots, K=K + 1, max_val=1e16, n_threads=4)
    with make_scan(n, h, 4, K, cx, rt, w, batch=8) as sc:
        rule = sc.tuning()["lig_pass"]
        for kw in ({"lig_pass": 1}, {"lig_pass": 2}, {"col_mz_extra": 1}, {"col_mz_extra": 2},
                   {"col_mz_extra": 15}):
            sc.set_tuning(**kw)
            eff = sc.tuning()
            for k, v in kw.items():
                assert eff[k] == v, (kw, eff)
            print(f"  {kw} (ligand pass rule {rule})", flush=True)
            check_grids(oracle, sc, n, h, cx.ligand.pos, w, S, R, rots[:2])
            peaks, nout = sc.run(rots)
            checked = compare_rotation_peaks(peaks, nout, val_o, off_o, s_o, tol, K)
            print(f"  checked {checked} of {len(rots) * K} ranks exactly", flush=True)
            assert checked >= len(rots) * K // 2


def test_tuning_errors():
    cx = synth.make_complex(300, 80, 45)
    rt, w = synth.terms_for(cx, 2)
    sc = pdfft.Scan((40, 40, 40), 1.5, 2, max_n_trans=4, max_val_corr=1e16, batch=8)
    sc.project_receptor(cx.receptor.pos, rt)
    with pytest.raises(pdfft.PdError) as e:
        sc.tuning()
    assert e.value.status == STATE
    sc.set_ligand(cx.ligand.pos, w)
    for kw in ({"col_groups": -1}, {"col_mz_extra": 16}, {"lig_pass": 3}, {"lanes": 4}):
        with pytest.raises(pdfft.PdError) as e:
            sc.set_tuning(**kw)
        assert e.value.status == INVALID, kw
    sc.close()
    # a split-plane grid has no whole-plane ligand pass
    sp = pdfft.Scan((32, 40, 40), 1.5, 2, max_n_trans=4, max_val_corr=1e16, batch=8, flags=pdfft.SCAN_SPLIT_PLANES)
    with pytest.raises(pdfft.PdError) as e:
        sp.set_tuning(lig_pass=1)
    assert e.value.status == UNSUPPORTED
    with pytest.raises(pdfft.PdError) as e:
        sp.set_tuning(lig_pass=2)
    assert e.value.status == UNSUPPORTED
    sp.close()

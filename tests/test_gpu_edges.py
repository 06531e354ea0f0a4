"""Edge cases of the scan on the GPU, each against the FP64 oracle or an exact
expectation:

  * no rotations: nothing to do, no error;
  * one rotation (a batch of one, the rotation-pair kernel's phantom partner)
    and one rotation per launch: the same bits as a full batch;
  * the largest maxNTrans (32): peaks against the oracle;
  * a one-atom ligand of unit weight: E(s) = R(l + s), the receptor grid read
    at the atom's cell shifted by s (docking_app.hpp:546-553);
  * a ligand of zero weights: every cell is exactly 0 (+0 or -0), so every
    cell ties; the reference's bounded heap keeps the first maxNTrans cells in
    row-major order (Common/queue.hpp:225-239, value < maxValCorr; equal values
    never replace), and with maxValCorr = 0 nothing passes.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from proddl2_amd import pdfft, synth  # noqa: E402

from tests.parity_helpers import (  # noqa: E402
    build_case, compare_rotation_peaks, grid_tolerance, make_scan, tolerance)

N48 = (48, 48, 48)


def test_zero_rotations():
    print("case zero rotations", flush=True)
    cx = synth.make_complex(300, 80, 61)
    rt, w = synth.terms_for(cx, 2)
    with make_scan(N48, 1.5, 2, 5, cx, rt, w) as sc:
        peaks, nout = sc.run(np.zeros((0, 9)))
        assert peaks.shape == (0, 5) and nout.shape == (0,)
        cand, ncand = sc.candidates(np.zeros((0, 9)))
        assert cand.shape[0] == 0 and ncand.shape == (0,)


# 160^3 runs the 1,024-thread whole-plane kernels (16 waves per block)
EDGE_CASES = [(N48, 1.5, 500, 150), ((160, 160, 160), 1.2, 2000, 500)]


@pytest.mark.parametrize("n,h,n_rec,n_lig", EDGE_CASES)
def test_one_rotation_and_one_per_launch_same_bits(oracle, n, h, n_rec, n_lig):
    print(f"case one rotation / batch 1, n={n}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, 2, seed=62)
    rots = synth.rotation_set(7, 63)
    with make_scan(n, h, 2, 6, cx, rt, w, batch=64) as sc:
        ref, nref = sc.run(rots)
        for i in range(len(rots)):
            p1, n1 = sc.run(rots[i:i + 1], rot_index0=i)
            assert n1[0] == nref[i] and np.array_equal(p1[0].view(np.uint8), ref[i].view(np.uint8)), i
        sc.set_batch(1)
        p, n = sc.run(rots)
        assert np.array_equal(n, nref) and np.array_equal(p.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("n,h,n_rec,n_lig", [(N48, 1.5, 700, 220), ((160, 160, 160), 1.2, 2000, 500)])
def test_largest_maxntrans_matches_oracle(oracle, n, h, n_rec, n_lig):
    """At 160^3 the inverse pass's block merge takes 16 waves x 32 keys, eight per lane."""
    K = 32
    print(f"case K={K}, n={n}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, 2, seed=64)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(5, 65)
    tol = tolerance(R, w, n)
    nout_o, off_o, s_o, val_o = oracle.scan(n, h, S, cx.ligand.pos, w, rots, K=K + 1, max_val=1e16,
                                            n_threads=4)
    with make_scan(n, h, 2, K, cx, rt, w, batch=4) as sc:
        peaks, nout = sc.run(rots)
    checked = compare_rotation_peaks(peaks, nout, val_o, off_o, s_o, tol, K)
    print(f"  checked {checked} of {len(rots) * K} ranks exactly", flush=True)
    assert checked >= len(rots) * K // 2


def test_one_atom_ligand_reads_the_receptor_grid(oracle):
    """T = 1, one ligand atom of weight 1: the correlation at displacement s is
    the receptor grid at the atom's cell + s, so the GPU grid equals the FP64
    receptor grid (rolled) within the FP32 bar, and the peaks are the
    receptor grid's lowest cells."""
    print("case one-atom ligand", flush=True)
    n, h = (40, 36, 44), 1.5
    cx = synth.make_complex(400, 20, 66)
    rt = synth.terms_ljcomp(cx, alpha=0.4, cutoff=9.0)[0][1:]  # the LJ attraction field alone
    lig = np.array([[0.37, -0.81, 1.26]])
    w = np.ones((1, 1))
    R = np.stack([oracle.project_receptor(n, h, cx.receptor.pos, k, p, cut, md) for (k, cut, md, p) in rt])
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(3, 67)
    K = 8
    nout_o, off_o, s_o, val_o = oracle.scan(n, h, S, lig, w, rots, K=K + 1, max_val=1e16, n_threads=2)
    tol = tolerance(R, w, n)
    with pdfft.Scan(n, h, 1, max_n_trans=K, max_val_corr=1e16, batch=4) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(lig, w)
        peaks, nout = sc.run(rots)
        for r in rots[:2]:
            q = oracle.rotate_points(r, lig)
            L = oracle.project_ligand(n, h, q, w[0])
            cell = np.argwhere(L != 0)[0]
            # C(x) = sum_b R(b) L(b + x) = R(cell - x) (periodic)
            want = np.roll(R[0][::-1, ::-1, ::-1], shift=tuple(1 + c for c in cell), axis=(0, 1, 2))
            got = sc.correlation_grid(r).astype(np.float64)
            err, bar = np.abs(got - want).max(), grid_tolerance(tol, want)
            print(f"  cell {cell}: max|err| {err:.3e} bar {bar:.3e}", flush=True)
            assert err <= bar
    checked = compare_rotation_peaks(peaks, nout, val_o, off_o, s_o, tol, K)
    assert checked >= len(rots) * K // 2


def test_zero_weights_tie_everywhere(oracle):
    print("case zero-weight ligand", flush=True)
    cx = synth.make_complex(300, 80, 68)
    rt, w = synth.terms_for(cx, 2)
    w0 = np.zeros_like(w)
    rots = synth.rotation_set(5, 69)
    K, k_inp = 6, 40
    with pdfft.Scan(N48, 1.5, 2, max_n_trans=K, max_n_trans_inp=k_inp, max_val_corr=1e16, batch=4) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w0)
        peaks, nout = sc.run(rots)
        assert (nout == K).all()
        assert (peaks["value"] == 0).all()
        assert (peaks["offset"] == np.arange(K)).all(), peaks["offset"]
        cand, ncand = sc.candidates(rots)
        assert (ncand == k_inp).all()
        assert (cand["offset"] == np.arange(k_inp)).all()
    # the oracle's bounded heap of maxNTransInp = maxNTrans cells keeps the same
    # cells.  (With a larger maxNTransInp the reference keeps the first
    # maxNTransInp cells and reports the first maxNTrans after sort_heap, whose
    # order among equal values is libstdc++'s: exact ties are the one place
    # where the order differs.)
    R = np.stack([oracle.project_receptor(N48, 1.5, cx.receptor.pos, k, p, cut, md) for (k, cut, md, p) in rt])
    _, off_o, _, _ = oracle.scan(N48, 1.5, oracle.receptor_spectra(R), cx.ligand.pos, w0, rots[:2], K=K,
                                 max_inp=K, max_val=1e16)
    assert all(sorted(o) == list(range(K)) for o in off_o)
    # maxValCorr = 0: no cell is below it
    with pdfft.Scan(N48, 1.5, 2, max_n_trans=K, max_val_corr=0.0, batch=4) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w0)
        peaks, nout = sc.run(rots)
        assert (nout == 0).all()


def test_dense_plane_beyond_the_atom_list(oracle):
    """A flat ligand of 900 atoms in one x plane (0.9 A apart in y and z, so
    some 1.2 A cells hold two atoms): more than the 512 atoms the 128^3 ligand
    pass lists per plane, so that plane's z spectra are summed from the atom
    cells directly.  Rotations about x keep the disk in its plane.  Grids
    within the bar of the oracle's, and the same bits at batch 1 as at batch 4."""
    from tests.parity_helpers import check_grids, oracle_receptor
    n, h, T = (128, 128, 128), 1.2, 4
    print(f"case dense ligand plane n={n}", flush=True)
    cx = synth.make_complex(1500, 60, 71)
    rt, _ = synth.terms_for(cx, T)
    g = np.arange(-15, 16) * 0.9
    yy, zz = np.meshgrid(g, g, indexing="ij")
    keep = yy ** 2 + zz ** 2 <= 13.6 ** 2
    disk = np.stack([np.full(keep.sum(), 0.3), yy[keep], zz[keep]], axis=1)
    assert disk.shape[0] > 512
    w = np.random.default_rng(72).uniform(-1.0, 1.0, (T, disk.shape[0]))
    R = oracle_receptor(oracle, n, h, cx.receptor.pos, rt)
    S = oracle.receptor_spectra(R)
    rots = []
    for a in (0.0, 0.7, 2.1):
        c, s = np.cos(a), np.sin(a)
        rots.append(np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64))
    rots = np.stack(rots)
    sc = pdfft.Scan(n, h, T, max_n_trans=10, max_val_corr=1e16, batch=4)
    with sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(disk, w)
        check_grids(oracle, sc, n, h, disk, w, S, R, rots)
        ref = sc.run(rots)
        sc.set_batch(1)
        got = sc.run(rots)
    assert np.array_equal(ref[0].view(np.uint8), got[0].view(np.uint8)) and np.array_equal(ref[1], got[1])
    print(f"  {disk.shape[0]} atoms in one plane: grids within the bar, batch 1 == batch 4", flush=True)


def test_six_terms_weights_beyond_the_list(oracle):
    """Six terms (the four-term set and the two-term set together): more terms
    than the ligand pass keeps weights of in its atom list, so it reads them
    from the weight array.  Grids within the bar and peaks against the oracle."""
    from tests.parity_helpers import check_grids, check_peaks_terms, oracle_receptor
    n, h = (48, 48, 48), 1.2
    print(f"case six terms n={n}", flush=True)
    cx = synth.make_complex(700, 220, 73)
    rt4, w4 = synth.terms_four(cx)
    rt2, w2 = synth.terms_shape_elec(cx)
    rt, w = rt4 + rt2, np.concatenate([w4, w2])
    R = oracle_receptor(oracle, n, h, cx.receptor.pos, rt)
    S = oracle.receptor_spectra(R)
    with make_scan(n, h, 6, 10, cx, rt, w) as sc:
        check_grids(oracle, sc, n, h, cx.ligand.pos, w, S, R, synth.rotation_set(2, seed=74))
    checked = check_peaks_terms(oracle, n, h, cx, rt, w, R, 0, n_rot=6)
    print(f"  {checked} ranks exact outside the band", flush=True)

"""GPU parity: libpdfft.so (through its C ABI) against the FP64 CPU oracle.

Bars:
  * receptor projection (FP64 on the GPU): bit-identical to the oracle;
  * correlation values (FP32 pipeline): |gpu - oracle| <= tol (see tolerance()); for
    whole grids the larger of tol and 4 eps32 max|C| (see grid_tolerance());
  * selected translations (integers): identical to the oracle's at every rank
    whose oracle value is separated from its neighbours by more than 2*tol;
  * FP64 rescoring: within 1e-9 relative of the oracle's direct sum.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # collected on CPU too, but every test needs the device
    pytest.skip("no GPU", allow_module_level=True)

from proddl2_amd import pdfft, synth  # noqa: E402

from tests.parity_helpers import (  # noqa: E402
    EPS32, build_case, check_grids, check_peaks_terms, grid_tolerance, make_scan, tolerance)


CASES = [
    # grid, h, n_rec, n_lig, terms.  Cubic edges in {32,48,64,96,120,128,144,160}
    # run the edge-specialised kernels, the rest the generic runtime-plan kernels.
    ((32, 32, 32), 1.5, 250, 60, 2),
    ((40, 36, 48), 1.5, 300, 80, 4),
    ((64, 64, 64), 1.2, 1800, 450, 2),
    ((96, 96, 96), 1.2, 1500, 300, 4),
    ((120, 120, 120), 1.2, 2000, 500, 2),
    # radix-7, -11 and -13 edges (findBestSize outputs such as 56, 44, 52; z half-length 26)
    ((56, 44, 52), 1.5, 300, 80, 2),
    # edge-specialised column pass (N0 = 64) behind the generic plane kernels, with the
    # ligand spectra in the column pass's interleaved order
    ((64, 48, 40), 1.5, 300, 80, 4),
    # planes over half of LDS: the ligand pass as row and column kernels, the
    # inverse as the whole-plane kernel
    ((144, 144, 144), 1.2, 2000, 500, 4),
    ((160, 160, 160), 1.2, 2000, 500, 2),
    # the same choice behind the runtime-plan kernels (150 x 76 plane: 92 KB of LDS)
    ((40, 150, 150), 1.5, 300, 80, 2),
    # odd last edges, as findBestSize returns them (the z transform is a full-length
    # complex FFT; the column pass runs the N0-specialised kernel at 64)
    ((45, 50, 75), 1.5, 300, 80, 2),
    ((77, 78, 91), 1.2, 1500, 300, 4),
    ((75, 75, 75), 1.2, 1500, 300, 2),
    ((64, 60, 81), 1.5, 700, 200, 4),
    # 135 x 135 plane (146 KB of LDS): ligand pass as row and column kernels
    ((135, 135, 135), 1.2, 2000, 500, 2),
]


@pytest.mark.parametrize("n,h,n_rec,n_lig,T", CASES)
def test_receptor_projection_bit_identical(oracle, n, h, n_rec, n_lig, T):
    print(f"case receptor projection n={n} T={T}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=11)
    with make_scan(n, h, T, 4, cx, rt, w) as sc:
        G = sc.receptor_grids()
    assert G.shape == R.shape
    bad = np.argwhere(G != R)
    assert bad.shape[0] == 0, f"{bad.shape[0]} cells differ, first {bad[:3]}: {G[tuple(bad[0])]} vs {R[tuple(bad[0])]}"


# The same cases through the split plane kernels (row and column passes),
# which large grids such as 256^3 use.
SPLIT = pdfft.SCAN_SPLIT_PLANES
CASES_BOTH = [c + (0,) for c in CASES] + [
    ((40, 36, 48), 1.5, 300, 80, 4, SPLIT),
    ((64, 64, 64), 1.2, 1800, 450, 2, SPLIT),
    ((96, 96, 96), 1.2, 1500, 300, 4, SPLIT),
    # z half-length 26 has no register-FFT split: the runtime-plan row kernel
    ((56, 44, 52), 1.5, 300, 80, 2, SPLIT),
    # odd last edges through the row and column kernels; 243 = 3^5 takes 16 rows per block
    ((40, 36, 45), 1.5, 300, 80, 4, SPLIT),
    ((32, 40, 243), 1.0, 300, 80, 2, SPLIT),
    # N1 = N2 = 256 (a 256 x 129 plane exceeds LDS): the ligand plane pass and the
    # inverse (y pass, rows, selection) each in one kernel, and the same case
    # through the separate row and column kernels
    ((32, 256, 256), 1.5, 300, 80, 2, 0),
    ((32, 256, 256), 1.5, 300, 80, 2, pdfft.SCAN_SPLIT_INVERSE | pdfft.SCAN_SPLIT_FORWARD),
]


@pytest.mark.parametrize("n,h,n_rec,n_lig,T,flags", CASES_BOTH)
def test_correlation_grid(oracle, n, h, n_rec, n_lig, T, flags):
    print(f"case correlation grid n={n} T={T} flags={flags}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=12)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(3, seed=3)
    tol = tolerance(R, w, n)
    with make_scan(n, h, T, 4, cx, rt, w, flags=flags) as sc:
        for r in rots:
            q = oracle.rotate_points(r, cx.ligand.pos)
            L = np.stack([oracle.project_ligand(n, h, q, w[t]) for t in range(T)])
            ref = oracle.correlate(S, L)
            got = sc.correlation_grid(r).astype(np.float64)
            err = np.abs(got - ref).max()
            bar = grid_tolerance(tol, ref)
            print(f"  max|err|={err:.3e} tol={tol:.3e} bar={bar:.3e} range=[{ref.min():.3e},{ref.max():.3e}]",
                  flush=True)
            assert err <= bar


# Ligand boxes of different widths select different column kernel variants:
# MZ, the number of x positions per lane that may hold ligand planes (the
# smallest of ceil(q A / 16), q = 1..16, covering the box: exact for A <= 16),
# and with it the zero-tail first DFT stage (MZ dividing A) or the full DFT.
COLUMN_VARIANTS = [
    ((32, 32, 32), 1.0, 200, 300, 2),       # A 4, MZ 4: full DFT
    ((64, 64, 64), 1.0, 400, 1500, 2),      # A 8, MZ 6: full DFT
    ((96, 96, 96), 1.0, 400, 3000, 2),      # A 12, MZ 7: full DFT
    ((128, 128, 128), 1.2, 1500, 1500, 4),  # A 16, MZ 5: full DFT
    ((128, 128, 128), 1.2, 1500, 800, 4),   # A 16, MZ 4: zero-tail DFT (R 4)
]


@pytest.mark.parametrize("n,h,n_rec,n_lig,T", COLUMN_VARIANTS)
def test_correlation_grid_column_variants(oracle, n, h, n_rec, n_lig, T):
    print(f"case column variant n={n} n_lig={n_lig}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=12)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(2, seed=5)
    tol = tolerance(R, w, n)
    with make_scan(n, h, T, 4, cx, rt, w) as sc:
        for r in rots:
            q = oracle.rotate_points(r, cx.ligand.pos)
            L = np.stack([oracle.project_ligand(n, h, q, w[t]) for t in range(T)])
            ref = oracle.correlate(S, L)
            got = sc.correlation_grid(r).astype(np.float64)
            err = np.abs(got - ref).max()
            print(f"  max|err|={err:.3e} tol={tol:.3e}", flush=True)
            assert err <= tol


def test_second_ligand_with_other_box(oracle):
    """A second set_ligand with a wider box moves the L1 plane origin, which the
    receptor spectra follow; the grids stay equal to the oracle's."""
    n, h, T = (64, 64, 64), 1.0, 2
    cx, rt, w, R = build_case(oracle, n, h, 400, 300, T, seed=12)
    big = synth.make_complex(400, 1500, 12)
    _, wb = synth.terms_for(big, T)
    S = oracle.receptor_spectra(R)
    r = synth.rotation_set(1, seed=6)[0]
    tol = tolerance(R, w, n)
    with make_scan(n, h, T, 4, cx, rt, w) as sc:
        for lig, wl in ((cx.ligand.pos, w), (big.ligand.pos, wb), (cx.ligand.pos, w)):
            sc.set_ligand(lig, wl)
            q = oracle.rotate_points(r, lig)
            L = np.stack([oracle.project_ligand(n, h, q, wl[t]) for t in range(T)])
            err = np.abs(sc.correlation_grid(r).astype(np.float64) - oracle.correlate(S, L)).max()
            print(f"  n_lig={len(lig)} max|err|={err:.3e} tol={tol:.3e}", flush=True)
            assert err <= tolerance(R, wl, n)


@pytest.mark.parametrize("n,h,n_rec,n_lig,T,flags", CASES_BOTH)
def test_scan_peaks_match_oracle(oracle, n, h, n_rec, n_lig, T, flags):
    print(f"case scan peaks n={n} T={T} flags={flags}", flush=True)
    check_peaks(oracle, n, h, n_rec, n_lig, T, flags, n_rot=12, seed=13)


def check_peaks(oracle, n, h, n_rec, n_lig, T, flags, n_rot, seed, K=10):
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=seed)
    return check_peaks_terms(oracle, n, h, cx, rt, w, R, flags, n_rot, K=K)


def test_one_kernel_inverse_same_bits_as_two(oracle):
    """N1 = N2 = 256: the one-kernel inverse (y pass, rows and selection with the
    plane on chip) does the arithmetic of the two inverse kernels in the same
    order: correlation grids and peaks are bit-identical."""
    n, h, T, K = (32, 256, 256), 1.5, 2, 10
    print(f"case one-kernel vs two-kernel inverse n={n}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, 300, 80, T, seed=14)
    rots = synth.rotation_set(9, seed=5)
    out = []
    for flags in (0, pdfft.SCAN_SPLIT_INVERSE):
        with make_scan(n, h, T, K, cx, rt, w, batch=4, flags=flags) as sc:
            grids = [sc.correlation_grid(r) for r in rots[:2]]
            peaks, nout = sc.run(rots)
        out.append((grids, peaks, nout))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b), np.abs(a - b).max()
    assert np.array_equal(out[0][2], out[1][2])
    for f in ("rot", "rank", "s", "offset", "value"):  # value64 is NaN (not rescored)
        assert np.array_equal(out[0][1][f], out[1][1][f]), f


def test_stress_256(oracle):
    """256^3 at 1.0 A, receptor 12,000 / ligand 3,000 atoms, 4 terms (the large-grid
    stress configuration): a 256 x 129 plane does not fit in LDS, so the scan runs
    the split plane kernels; receptor grids bit-identical, peaks against the oracle."""
    print("case stress 256^3", flush=True)
    check_peaks(oracle, (256, 256, 256), 1.0, 12000, 3000, 4, 0, n_rot=2, seed=17)


def test_rescore64_matches_direct_sum(oracle):
    n, h, T = (40, 36, 48), 1.5, 4
    cx, rt, w, R = build_case(oracle, n, h, 300, 80, T, seed=14)
    rots = synth.rotation_set(6, seed=5)
    rng = np.random.default_rng(0)
    s = rng.integers(-10, 10, size=(6, 3)).astype(np.int32)
    with make_scan(n, h, T, 4, cx, rt, w) as sc:
        e = sc.rescore64(rots, s)
    for i in range(6):
        q = oracle.rotate_points(rots[i], cx.ligand.pos)
        ref = oracle.direct_energy(R, oracle.cells(n, h, q), w, s[i])
        assert abs(e[i] - ref) <= 1e-9 * max(1.0, abs(ref)), (i, e[i], ref)


def test_max_val_corr_filter(oracle):
    n, h, T = (32, 32, 32), 1.5, 2
    cx, rt, w, R = build_case(oracle, n, h, 250, 60, T, seed=15)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(4, seed=6)
    nout_o, off_o, s_o, val_o = oracle.scan(n, h, S, cx.ligand.pos, w, rots, K=10, max_val=1e16)
    tol = tolerance(R, w, n)
    thr = float(np.median(val_o[:, 5]))
    with make_scan(n, h, T, 10, cx, rt, w, max_val=thr) as sc:
        peaks, nout = sc.run(rots)
    exact = 0
    for r in range(len(rots)):
        assert (peaks["value"][r, : nout[r]] < thr).all()
        expect = int((val_o[r] < thr).sum())
        # Only a value within the FP32 band of the threshold may fall on the
        # other side of it on the GPU: the count may differ by at most the number
        # of such values (the threshold is an oracle value, so each rotation
        # whose 6th value set it has at least one).
        ambiguous = int((np.abs(val_o[r] - thr) <= tol).sum())
        assert abs(int(nout[r]) - expect) <= ambiguous, (r, nout[r], expect, ambiguous)
        exact += ambiguous == 0
        # the kept values are the oracle's first nout values, within the band
        assert np.all(np.abs(peaks["value"][r, : nout[r]] - val_o[r, : nout[r]]) <= 3 * tol)
    print(f"  {exact} of {len(rots)} rotations had no value within the band of maxValCorr", flush=True)


def test_ligand_outside_grid_is_domain_error(oracle):
    n, h, T = (32, 32, 32), 1.5, 2
    cx, rt, w, R = build_case(oracle, n, h, 250, 60, T, seed=16)
    with pdfft.Scan(n, h, T, max_n_trans=4, max_val_corr=1e16) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        far = cx.ligand.pos * 3.0  # radius ~ 22 A > half box 24 A after rotation for some atoms
        far[0] = [30.0, 0.0, 0.0]
        sc.set_ligand(far, w)
        with pytest.raises(pdfft.PdError) as ei:
            sc.run(synth.rotation_set(2, seed=1))
        assert ei.value.status == 2


def test_call_order_is_state_error():
    with pdfft.Scan((32, 32, 32), 1.5, 2, max_n_trans=4) as sc:
        with pytest.raises(pdfft.PdError) as ei:
            sc.run(synth.rotation_set(1, seed=1))
        assert ei.value.status == 6


def test_every_find_best_size_edge_is_accepted():
    """Every findBestSize edge up to 256 (Math/fft.hpp:141-168: {2,3,5,7,11,13}-smooth
    with e11 + e13 <= 1, odd ones included) is accepted on each axis."""
    edges = sorted({pdfft.find_best_size(k) for k in range(8, 257)} - {257})
    edges = [e for e in edges if e <= 256]
    assert 75 in edges and 91 in edges and 135 in edges and 243 in edges
    for e in edges:
        for n in ((e, 32, 32), (32, e, 32), (32, 32, e)):
            with pdfft.Scan(n, 1.5, 2, max_n_trans=4, batch=4):
                pass
    print(f"  {len(edges)} edges accepted on every axis", flush=True)

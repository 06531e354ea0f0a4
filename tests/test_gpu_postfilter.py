"""Many translations per rotation and the post-filters on the GPU.

(a) pd_scan_candidates (CorrelationProcessor::selectFromFFT with maxNTransInp,
    docking_fft.hpp:316-345) against the oracle's bounded-queue selection on its
    FP64 grids: counts equal, offsets exact at every rank outside twice the
    measured grid error, values within it.
(b) TranSymm / TranClust (docking_fft.hpp:30-230) run on the GPU against the
    oracle's restatements applied to the GPU's own candidate list: identical
    output (rotations whose candidates hold exactly equal values, whose
    TranClust order the reference leaves to std::sort, are counted apart), on
    a whole-plane grid, an odd last edge and a split-plane grid.
(c) A graph with the post-filters replays identically to the eager scan.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

import torch  # noqa: E402

from proddl2_amd import pdfft, synth  # noqa: E402

from tests.parity_helpers import build_case  # noqa: E402


def _axis_rot(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def _rotations(n, seed):
    """Half two-fold rotations (so that some poses pass a C2 test), half random."""
    rng = np.random.default_rng(seed)
    R = [_axis_rot(rng.normal(size=3), np.pi) for _ in range(n // 2)]
    R += list(synth.random_rotation(rng, n - n // 2))
    return np.stack([r.reshape(9) for r in R])


def _rigid(rng):
    m = synth.random_rotation(rng)
    return np.concatenate([m.reshape(9), rng.normal(0, 3, 3)])


CAND_CASES = [
    # grid, h, n_rec, n_lig, terms, k_inp, max_val (None: the median of the best values)
    ((48, 48, 48), 1.5, 700, 220, 2, 300, 1e16),
    ((40, 36, 45), 1.5, 300, 80, 4, 1000, 1e16),
    ((64, 64, 64), 1.2, 1800, 450, 2, 4096, None),
    ((32, 256, 256), 1.5, 300, 80, 2, 500, 1e16),  # one-kernel split inverse writes the grid
]


@pytest.mark.parametrize("n,h,n_rec,n_lig,T,k_inp,max_val", CAND_CASES)
def test_candidates_match_oracle_selection(oracle, n, h, n_rec, n_lig, T, k_inp, max_val):
    print(f"case candidates n={n} k_inp={k_inp} max_val={max_val}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=31)
    S = oracle.receptor_spectra(R)
    rots = _rotations(6, 5)
    nout_o, off_o, s_o, val_o, grids = oracle.scan(n, h, S, cx.ligand.pos, w, rots, K=1, max_val=1e16,
                                                   n_threads=4, want_grids=True)
    if max_val is None:
        max_val = float(np.median([np.sort(gr.reshape(-1))[k_inp // 2] for gr in grids]))
    with pdfft.Scan(n, h, T, max_n_trans=10, max_n_trans_inp=k_inp, max_val_corr=max_val, batch=4) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w)
        E = np.array([np.abs(sc.correlation_grid(r).astype(np.float64) - gr).max() for r, gr in zip(rots, grids)])
        cand, ncand = sc.candidates(rots, rot_index0=7)
    exact = 0
    for r in range(len(rots)):
        o_off, o_s, _, o_val = oracle.select(grids[r], h, k_inp, max_val, k_inp)
        band = 2 * E[r]
        m = int(ncand[r])
        # the count differs only by values within the band of maxValCorr
        amb = int((np.abs(grids[r] - max_val) <= E[r]).sum())
        assert abs(m - len(o_off)) <= amb, (r, m, len(o_off), amb)
        c = cand[r, :m]
        assert (c["rot"] == 7 + r).all() and (c["rank"] == np.arange(m)).all()
        assert np.all(np.diff(c["value"]) >= 0)
        k = min(m, len(o_off))
        v = np.concatenate([o_val, [np.inf]])
        for i in range(k):
            lo = v[i] - v[i - 1] if i else np.inf
            if lo > band and v[i + 1] - v[i] > band:
                assert c["offset"][i] == o_off[i], (r, i)
                assert (c["s"][i] == o_s[i]).all()
                exact += 1
            assert abs(c["value"][i] - v[i]) <= E[r] + 1e-12 * abs(v[i])
    print(f"  {exact} of {int(ncand.sum())} candidates exact outside the measured band (E max {E.max():.2e})",
          flush=True)
    assert exact >= 0.5 * ncand.sum()


def _symm_keep(oracle, c, rot, h, pf):
    tran = h * c["s"].astype(np.float64)
    keep, rmsd = oracle.tran_symm(tran, c["value"].astype(np.float64), rot, pf["frame_rec_inv"],
                                  pf["frame_lig"], pf["n_multimer"], pf["ligand_radius"], pf["max_rmsd_symm"])
    return keep, rmsd


POST_GRIDS = [((48, 48, 48), 700, 220), ((45, 50, 75), 300, 80), ((32, 256, 256), 300, 80)]


@pytest.mark.parametrize("mode", ["symm", "clust", "both"])
@pytest.mark.parametrize("n,n_rec,n_lig", POST_GRIDS)
def test_postfilters_match_oracle_on_gpu_candidates(oracle, mode, n, n_rec, n_lig):
    h, T, K, k_inp = 1.5, 2, 10, 400
    print(f"case post-filters {mode}: n={n} K={K} k_inp={k_inp}", flush=True)
    cx, rt, w, R = build_case(oracle, n, h, n_rec, n_lig, T, seed=32)
    rng = np.random.default_rng(9)
    rots = _rotations(16, 6)
    pf = dict(n_multimer=2 if mode != "clust" else 0, max_rmsd_symm=8.0, ligand_radius=9.0,
              frame_rec_inv=_rigid(rng), frame_lig=_rigid(rng), cluster=mode != "symm", cluster_radius=4.0)
    if mode != "clust":
        # a conjugating pair of frames keeps two-fold poses two-fold
        fb = pf["frame_lig"]
        m = fb[:9].reshape(3, 3)
        pf["frame_rec_inv"] = np.concatenate([m.T.reshape(9), -m.T @ fb[9:]])
    with pdfft.Scan(n, h, T, max_n_trans=K, max_n_trans_inp=k_inp, max_val_corr=1e16, batch=5) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w)
        cand, ncand = sc.candidates(rots)
        sc.set_postfilter(**pf)
        peaks, nout = sc.run(rots)
        sc.set_postfilter()  # cleared: the plain scan again
        plain, _ = sc.run(rots)
    assert np.array_equal(plain["offset"], cand["offset"][:, :K])
    tied, exact, kept_symm = 0, 0, 0
    for r in range(len(rots)):
        c = cand[r, : ncand[r]]
        idx = np.arange(len(c))
        if pf["n_multimer"] > 1:
            keep, rmsd = _symm_keep(oracle, c, rots[r].reshape(3, 3), h, pf)
            kept_symm += len(keep)
            idx = idx[keep]
        if pf["cluster"]:
            vals = c["value"][idx].astype(np.float64)
            if len(np.unique(vals)) < len(vals):
                tied += 1
                continue
            lead = oracle.tran_clust(h * c["s"][idx].astype(np.float64), vals, pf["cluster_radius"], K,
                                     cap=k_inp)
            idx = idx[lead]
        want = idx[:K]
        got = peaks[r, : nout[r]]
        assert nout[r] == len(want), (r, nout[r], len(want))
        assert np.array_equal(got["offset"], c["offset"][want]), r
        assert np.array_equal(got["value"], c["value"][want]) and (got["rank"] == np.arange(len(want))).all()
        exact += 1
    print(f"  {exact} rotations identical, {tied} with exactly equal candidate values skipped; "
          f"{kept_symm} candidates passed the symmetry test", flush=True)
    assert exact >= len(rots) - 2
    if pf["n_multimer"] > 1:
        assert 0 < kept_symm < ncand.sum()


def test_graph_with_postfilters_matches_eager(oracle):
    dev = torch.device("cuda:0")
    n, h, T, K = (48, 48, 48), 1.5, 2, 8
    cx, rt, w, R = build_case(oracle, n, h, 700, 220, T, seed=33)
    rots = _rotations(40, 7)
    with pdfft.Scan(n, h, T, max_n_trans=K, max_n_trans_inp=500, max_val_corr=1e16, batch=8) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w)
        sc.set_postfilter(n_multimer=2, ligand_radius=9.0, cluster=True, cluster_radius=3.0)
        ref, nref = sc.run(rots)
        d_r = torch.from_numpy(rots).to(dev)
        d_p = torch.zeros((len(rots), K * pdfft.PEAK_DTYPE.itemsize // 4), dtype=torch.int32, device=dev)
        d_n = torch.zeros(len(rots), dtype=torch.int32, device=dev)
        g = pdfft.Graph([(sc, len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0)], 1)
        for _ in range(2):
            g.launch()
            g.sync()
            torch.cuda.synchronize()
            got = d_p.cpu().numpy().view(pdfft.PEAK_DTYPE).reshape(len(rots), K)
            assert np.array_equal(d_n.cpu().numpy(), nref)
            for r in range(len(rots)):
                assert np.array_equal(got[r, : nref[r]].view(np.uint8), ref[r, : nref[r]].view(np.uint8))
            d_p.zero_()
        g.close()

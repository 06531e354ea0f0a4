"""GPU tests of the learned rotation proposal (libpdfft.so, include/pdprop.h)
against the CPU oracle (oracle/proposal_oracle.py).

Tolerances: maps and rotation functions are FP32 sums (LDS atomics, any
order) of up to ~1e4 FP64 terms, so they match the FP64 oracle to a few
1e-6 relative (sqrt(n) * 2^-24).  The
network runs its wide products on MFMA with BF16 operands; the oracle
rounds the same operands to BF16 and accumulates in FP64, so the remaining
difference comes from FP32 accumulation, and from the rare operand whose
FP32 and FP64 values round to neighbouring BF16 numbers (one BF16 ulp,
2^-8 relative, on that element).
"""
import os

import numpy as np
import pytest

from oracle import proposal_oracle as po
from proddl2_amd import pdfft, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def proposal(dev):
    from proddl2_amd.proposal import Proposal
    p = Proposal(max_rot=30000, seed=3)
    yield p
    p.close()


def _small_complex(seed=21):
    return synth.make_complex(500, 160, seed)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(scale=0.15, size=po.NPARAM)).astype(np.float32)


def test_maps_match_oracle(proposal):
    cx = _small_complex()
    proposal.set_complex(cx.receptor.pos, cx.ligand.pos)
    rec, lig = proposal.maps()
    for got, pos in ((rec, cx.receptor.pos), (lig, cx.ligand.pos)):
        want = po.maps(pos)
        scale = np.abs(want).max()
        assert scale > 0
        # every cell is an FP32 sum of up to ~1e4 positive FP64 terms
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6 * scale)


def _predict(proposal, rots, dev):
    d_r = torch.from_numpy(np.ascontiguousarray(rots, np.float64)).to(dev)
    d_s = torch.empty(rots.shape[0], dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    proposal.predict_device(rots.shape[0], d_r.data_ptr(), d_s.data_ptr())
    proposal.sync()
    return d_s.cpu().numpy()


def test_predict_matches_oracle(proposal, dev):
    cx = _small_complex()
    proposal.set_complex(cx.receptor.pos, cx.ligand.pos)
    params = _random_params(4)
    proposal.set_params(params)
    rec, lig = proposal.maps()
    rots = synth.rotation_set(96, 5)
    got = _predict(proposal, rots, dev)
    want, _ = po.forward(params.astype(np.float64), rec.astype(np.float64), lig.astype(np.float64), rots, bf=True)
    scale = max(1.0, np.abs(want).max())
    assert np.abs(got - want).max() <= 2e-3 * scale, np.abs(got - want).max()


def test_gradients_match_oracle(proposal):
    cx = _small_complex(22)
    proposal.set_complex(cx.receptor.pos, cx.ligand.pos)
    params = _random_params(6)
    proposal.set_params(params)
    rec, lig = proposal.maps()
    rots = synth.rotation_set(80, 7)
    energy = np.random.default_rng(8).normal(size=80).astype(np.float32)
    g, loss = proposal.gradients(rots, energy)
    wl, wg, _ = po.loss_and_grad(params.astype(np.float64), rec.astype(np.float64), lig.astype(np.float64), rots,
                                 energy.astype(np.float64), bf=True)
    assert abs(loss - wl) <= 1e-4 * abs(wl)
    blocks = [("conv_rec", slice(0, po.CONV_STACK)), ("conv_lig", slice(po.CONV_STACK, po.OFF_W1)),
              ("W1", slice(po.OFF_W1, po.OFF_B1)), ("b1", slice(po.OFF_B1, po.OFF_W2)),
              ("W2", slice(po.OFF_W2, po.OFF_B2)), ("b2", slice(po.OFF_B2, po.OFF_W3)),
              ("W3b3", slice(po.OFF_W3, po.NPARAM))]
    for name, sl in blocks:
        err = np.linalg.norm(g[sl] - wg[sl]) / np.linalg.norm(wg[sl])
        assert err <= 5e-3, (name, err)


def test_select_matches_oracle(proposal, dev):
    rng = np.random.default_rng(9)
    for n, keep in ((30000, 4500), (1000, 1), (777, 777), (5000, 1234)):
        s = np.round(rng.normal(size=n), 2).astype(np.float32)  # many exact ties
        d_s = torch.from_numpy(s).to(dev)
        d_i = torch.full((keep,), -1, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        proposal.select_device(n, d_s.data_ptr(), keep, d_i.data_ptr())
        proposal.sync()
        assert np.array_equal(d_i.cpu().numpy(), po.select_lowest(s, keep)), (n, keep)


def test_pruned_scan_reproduces_kept_rotations(proposal, dev):
    from proddl2_amd.proposal import pruned_buffers, run_pruned
    cx = synth.make_complex(600, 200, 31)
    rt, w = synth.terms_for(cx, 2)
    rots = synth.rotation_set(300, 32)
    with pdfft.Scan((48, 48, 48), 1.2, 2, max_n_trans=5, max_val_corr=1e16) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w)
        full, nfull = sc.run(rots)
        proposal.set_stream(sc.stream())
        try:
            proposal.set_complex(cx.receptor.pos, cx.ligand.pos)
            d_rots = torch.from_numpy(rots).to(dev)
            torch.cuda.synchronize()
            bufs = pruned_buffers(sc, proposal, d_rots, 45)
            run_pruned(sc, proposal, d_rots, bufs, 45)
            sc.sync()
        finally:
            proposal.set_stream(None)
        idx = bufs[1].cpu().numpy()
        peaks = bufs[3].cpu().numpy().view(pdfft.PEAK_DTYPE).reshape(45, 5)
        assert np.all(np.diff(idx) > 0)
        for j, r in enumerate(idx):
            assert np.array_equal(peaks[j]["rot"], full[r]["rot"])
            assert np.array_equal(peaks[j]["offset"], full[r]["offset"])
            assert np.array_equal(peaks[j]["value"], full[r]["value"])


def test_training_steps_reduce_loss(dev):
    from proddl2_amd.proposal import Proposal
    cx = _small_complex(41)
    rots = synth.rotation_set(512, 42)
    # a smooth synthetic target over the rotations of this complex
    energy = (np.cos(3 * rots[:, 0]) + rots[:, 4] * rots[:, 8]).astype(np.float32)
    with Proposal(max_rot=512, seed=1) as p:
        p.set_complex(cx.receptor.pos, cx.ligand.pos)
        d_r = torch.from_numpy(rots).to(dev)
        d_e = torch.from_numpy(energy).to(dev)
        torch.cuda.synchronize()
        losses = [p.train_step_device(512, d_r.data_ptr(), d_e.data_ptr(), 1e-3) for _ in range(60)]
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.3 * losses[0], losses[::10]


def test_fixed_score_params_rank_by_rotation_function_sum(proposal, dev):
    from proddl2_amd.proposal import fixed_score_params
    cx = _small_complex(51)
    proposal.set_complex(cx.receptor.pos, cx.ligand.pos)
    p = fixed_score_params()
    proposal.set_params(p)
    rec, lig = proposal.maps()
    rots = synth.rotation_set(200, 52)
    got = _predict(proposal, rots, dev)
    _, k = po.forward(p.astype(np.float64), rec.astype(np.float64), lig.astype(np.float64), rots, bf=False)
    want = -k["X0"][:, :po.NF].sum(1)
    # the 24 standardised inputs pass through BF16 (2^-9 relative each), and so
    # does their sum in the hidden layer
    assert np.abs(got - want).max() <= 2 * 24 * 2 ** -9 * np.abs(k["X0"][:, :po.NF]).max()


WEIGHTS_PRESENT = os.path.exists(os.path.join(os.path.dirname(pdfft.__file__), "data", "proposal_weights.npy"))


@pytest.mark.skipif(not WEIGHTS_PRESENT, reason="no trained weights")
def test_trained_network_recall_and_gain_at_128(proposal):
    """128^3 x 4 terms, 30,000 rotations, complexes never seen in training:
    recall of the exhaustive top-100 poses when 15% / 10% of the rotations are
    scanned, for the trained network and for the same pipeline without
    learning (identity convolutions, score = -sum of rotation functions)."""
    from proddl2_amd.proposal import evaluate, fixed_score_params, WEIGHTS
    quiet = lambda m: None  # noqa: E731
    for seed in (2024, 12):
        proposal.load(WEIGHTS)
        trained = evaluate(proposal, seed=seed, keep_fracs=(0.10, 0.15), repeats=1, log=quiet)
        proposal.set_params(fixed_score_params())
        fixed = evaluate(proposal, seed=seed, keep_fracs=(0.10,), repeats=1, log=quiet)
        r10, r15 = trained["keep"][0]["recall"], trained["keep"][1]["recall"]
        assert r15 >= 0.95 and r10 >= 0.95, trained
        assert r10 >= fixed["keep"][0]["recall"], (trained, fixed)
        print(f"seed {seed}: recall@100 trained {r10:.2f} (10%) {r15:.2f} (15%), "
              f"without learning {fixed['keep'][0]['recall']:.2f} (10%)")

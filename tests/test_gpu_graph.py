"""Several scans in one HIP graph (pd_graph_*) and batch-size changes
(pd_scan_set_batch): results must be bit-identical to eager scans.

The graph is built node by node (kernel nodes with explicit edges, no stream
capture).  Every test replays a graph several times with the program's own
eager work and synchronisations between the replays, and checks every replay."""
import numpy as np
import pytest

from proddl2_amd import pdfft, synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _pair(dev, n, n_rec, n_lig, seed, n_rot, terms=2, batch=64):
    cx = synth.make_complex(n_rec, n_lig, seed)
    rt, w = synth.terms_for(cx, terms)
    sc = pdfft.Scan((n, n, n), 1.2, terms, max_n_trans=6, max_val_corr=1e16, batch=batch)
    sc.project_receptor(cx.receptor.pos, rt)
    sc.set_ligand(cx.ligand.pos, w)
    rots = synth.rotation_set(n_rot, seed + 1)
    d_r = torch.from_numpy(rots).to(dev)
    d_p = torch.zeros((n_rot, 6 * pdfft.PEAK_DTYPE.itemsize // 4), dtype=torch.int32, device=dev)
    d_n = torch.zeros(n_rot, dtype=torch.int32, device=dev)
    return sc, rots, d_r, d_p, d_n


def _peaks(d_p, d_n, n_rot):
    torch.cuda.synchronize()
    return d_p.cpu().numpy().copy(), d_n.cpu().numpy().copy()


def test_batch_size_does_not_change_results(dev):
    sc, rots, d_r, d_p, d_n = _pair(dev, 48, 700, 220, 61, 150)
    ref, _ = sc.run(rots)
    for b in (1, 16, 37, 150):
        sc.set_batch(b)
        got, _ = sc.run(rots)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), b
    sc.close()


@pytest.mark.skip(reason="written for the stream-capture pd_graph_create, whose two-lane capture crashed the host process (cause not shown: the crashing code was never committed and its log is lost); capture was removed, graphs are now built node by node and tested by the test_built_graph_* tests")
def test_graph_replays_match_eager(dev):
    pairs = [_pair(dev, 48, 700, 220, 71, 200), _pair(dev, 64, 1500, 400, 72, 130, terms=4, batch=32),
             _pair(dev, 48, 500, 150, 73, 90, batch=16)]
    eager = []
    for sc, rots, d_r, d_p, d_n in pairs:
        sc.run_device(len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0, sync=True)
        eager.append(_peaks(d_p, d_n, len(rots)))
        d_p.zero_()
        d_n.zero_()
    torch.cuda.synchronize()
    for lanes, lane_of in ((1, [0, 0, 0]), (2, [0, 1, 0])):
        g = pdfft.Graph([(sc, len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), lane_of[i])
                         for i, (sc, rots, d_r, d_p, d_n) in enumerate(pairs)], lanes)
        for replay in range(3):
            g.launch()
            g.sync()
            for (sc, rots, d_r, d_p, d_n), (ep, en) in zip(pairs, eager):
                gp, gn = _peaks(d_p, d_n, len(rots))
                assert np.array_equal(gn, en)
                assert np.array_equal(gp, ep), (lanes, replay)
            # the program's own eager work between replays
            sc0, rots0, d_r0, d_p0, d_n0 = pairs[0]
            sc0.run_device(len(rots0), d_r0.data_ptr(), d_p0.data_ptr(), d_n0.data_ptr(), 0, sync=True)
            for _, _, _, d_p, d_n in pairs:
                d_p.zero_()
                d_n.zero_()
            torch.cuda.synchronize()
        g.close()
    for p in pairs:
        p[0].close()


@pytest.mark.skip(reason="written for the stream-capture pd_graph_create, whose two-lane capture crashed the host process (cause not shown: the crashing code was never committed and its log is lost); capture was removed, graphs are now built node by node and tested by the test_built_graph_* tests")
def test_graph_reports_domain_error(dev):
    sc, rots, d_r, d_p, d_n = _pair(dev, 48, 700, 220, 81, 40)
    # move one rotation's ligand far outside the grid by scaling the matrix
    bad = rots.copy()
    bad[5] *= 40.0
    d_bad = torch.from_numpy(bad).to(dev)
    g = pdfft.Graph([(sc, len(rots), d_bad.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0)], 1)
    g.launch()
    with pytest.raises(pdfft.PdError) as e:
        g.sync()
    assert e.value.status == 2
    g.close()
    # the handle is usable again afterwards
    got, _ = sc.run(rots)
    assert got.shape == (40, 6)
    sc.close()


def _eager(pairs):
    out = []
    for sc, rots, d_r, d_p, d_n in pairs:
        sc.run_device(len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0, sync=True)
        out.append(_peaks(d_p, d_n, len(rots)))
        d_p.zero_()
        d_n.zero_()
    torch.cuda.synchronize()
    return out


def test_built_graph_replays_match_eager(dev):
    print("case built graph: 3 pairs, lanes 1 and 2", flush=True)
    pairs = [_pair(dev, 48, 700, 220, 71, 200), _pair(dev, 64, 1500, 400, 72, 130, terms=4, batch=32),
             _pair(dev, 48, 500, 150, 73, 90, batch=16)]
    eager = _eager(pairs)
    for lanes, lane_of in ((1, [0, 0, 0]), (2, [0, 1, 0]), (3, [2, 1, 0])):
        g = pdfft.Graph([(sc, len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), lane_of[i])
                         for i, (sc, rots, d_r, d_p, d_n) in enumerate(pairs)], lanes)
        for replay in range(3):
            g.launch()
            g.sync()
            for (sc, rots, d_r, d_p, d_n), (ep, en) in zip(pairs, eager):
                gp, gn = _peaks(d_p, d_n, len(rots))
                assert np.array_equal(gn, en)
                assert np.array_equal(gp, ep), (lanes, replay)
            # the program's own eager work between replays
            sc0, rots0, d_r0, d_p0, d_n0 = pairs[0]
            sc0.run_device(len(rots0), d_r0.data_ptr(), d_p0.data_ptr(), d_n0.data_ptr(), 0, sync=True)
            for _, _, _, d_p, d_n in pairs:
                d_p.zero_()
                d_n.zero_()
            torch.cuda.synchronize()
        g.close()
        print(f"  lanes {lanes}: 3 replays identical to eager", flush=True)
    for p in pairs:
        p[0].close()


def test_built_graph_reports_domain_error(dev):
    sc, rots, d_r, d_p, d_n = _pair(dev, 48, 700, 220, 81, 40)
    bad = rots.copy()
    bad[5] *= 40.0  # one rotation's ligand far outside the grid
    d_bad = torch.from_numpy(bad).to(dev)
    g = pdfft.Graph([(sc, len(rots), d_bad.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0)], 1)
    g.launch()
    with pytest.raises(pdfft.PdError) as e:
        g.sync()
    assert e.value.status == 2
    # the flag was cleared: a replay with good rotations in the same buffers passes
    d_bad.copy_(d_r)
    g.launch()
    g.sync()
    g.close()
    got, _ = sc.run(rots)
    assert got.shape == (40, 6)
    sc.close()


def test_built_graph_holds_its_scans(dev):
    """While a graph names a scan's buffers they cannot move: set_batch,
    set_ligand and destroy return PD_ERR_STATE; after the graph is gone they work."""
    sc, rots, d_r, d_p, d_n = _pair(dev, 48, 700, 220, 91, 70)
    g = pdfft.Graph([(sc, len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0)], 1)
    with pytest.raises(pdfft.PdError) as e:
        g2 = pdfft.Graph([(sc, len(rots), d_r.data_ptr(), d_p.data_ptr(), d_n.data_ptr(), 0)], 1)
    assert e.value.status == 6
    for call in (lambda: sc.set_batch(16), lambda: sc.close()):
        with pytest.raises(pdfft.PdError) as e:
            call()
        assert e.value.status == 6
    cx = synth.make_complex(700, 220, 91)
    _, w = synth.terms_for(cx, 2)
    with pytest.raises(pdfft.PdError) as e:
        sc.set_ligand(cx.ligand.pos, w)
    assert e.value.status == 6
    g.launch()
    g.sync()
    ref = _peaks(d_p, d_n, len(rots))
    g.close()
    sc.set_batch(16)
    got, nout = sc.run(rots)
    assert np.array_equal(got.view(np.uint8).reshape(len(rots), -1),
                          ref[0].view(np.uint8).reshape(len(rots), -1))
    sc.close()


def test_built_graph_uneven_pairs(dev, oracle):
    """The batched uneven workload (32 pairs, receptor 800-6,000 / ligand
    200-2,000 atoms, grid edges from {64, ..., 160}, 4 terms) with 256
    rotations per pair, as one graph on 2 lanes: every replay byte-identical to
    the eager scans, and for each grid edge one pair's peaks against the oracle."""
    from proddl2_amd import batch
    from tests.parity_helpers import compare_rotation_peaks, oracle_receptor, tolerance
    specs = batch.uneven_pairs(32, 256, 3000)
    pairs = [batch.Pair(s, torch, dev) for s in specs]
    print(f"case uneven pairs: {len(pairs)} pairs, edges {sorted({s['edge'] for s in specs})}", flush=True)
    for p in pairs:
        p.run()
    eager = [p.peaks() for p in pairs]
    for p in pairs:
        p.d_peaks.zero_()
        p.d_nout.zero_()
    torch.cuda.synchronize()
    g = batch.make_graph(pairs, 2)
    for replay in range(2):
        g.launch()
        g.sync()
        for p, (ep, en) in zip(pairs, eager):
            gp, gn = p.peaks()
            assert np.array_equal(gn, en)
            assert np.array_equal(gp.view(np.uint8), ep.view(np.uint8)), (p.spec, replay)
        pairs[0].run(64)
        for p in pairs:
            p.d_peaks.zero_()
        torch.cuda.synchronize()
    g.close()
    print("  2 replays identical to eager", flush=True)
    seen = set()
    for p, (ep, en) in zip(pairs, eager):
        e = p.spec["edge"]
        if e in seen:
            continue
        seen.add(e)
        cx = synth.make_complex(p.spec["n_rec"], p.spec["n_lig"], p.spec["seed"])
        rt, w = synth.terms_for(cx, batch.N_TERMS)
        n = (e, e, e)
        R = oracle_receptor(oracle, n, p.spec["h"], cx.receptor.pos, rt)
        S = oracle.receptor_spectra(R)
        rots = synth.rotation_set(p.spec["n_rot"], p.spec["rot_seed"])[:2]
        nout_o, off_o, s_o, val_o = oracle.scan(n, p.spec["h"], S, cx.ligand.pos, w, rots, K=batch.K + 1,
                                                max_val=1e16, n_threads=2)
        tol = tolerance(R, w, n)
        checked = compare_rotation_peaks(ep[:2], en[:2], val_o, off_o, s_o, tol, batch.K)
        print(f"  edge {e}: {checked} of {2 * batch.K} ranks exact against the oracle", flush=True)
        assert checked >= batch.K
    for p in pairs:
        p.close()


def test_autotune_keeps_bits_and_persists(dev, tmp_path, monkeypatch):
    """batch.autotune on three uneven pairs: every tuned pair's peaks equal its
    untuned peaks byte for byte, the choices land in the tuning file, and a
    second call takes them from there without timing."""
    from proddl2_amd import batch
    path = str(tmp_path / "tuning.json")
    monkeypatch.setattr(batch, "TUNING_FILE", path)
    specs = batch.uneven_pairs(3, 192, 3100)
    pairs = [batch.Pair(s, torch, dev) for s in specs]
    print(f"case autotune: edges {[s['edge'] for s in specs]}", flush=True)
    untuned = []
    for p in pairs:
        p.run()
        untuned.append(p.peaks())
    logs = []
    tuned, source, _ = batch.autotune(pairs, torch, lanes=(), sample=128, log=logs.append)
    assert set(source.values()) == {"tuned"}
    for p, (pk, no) in zip(pairs, untuned):
        p.d_peaks.zero_()
        p.run()
        gp, gn = p.peaks()
        assert np.array_equal(gn, no) and np.array_equal(gp.view(np.uint8), pk.view(np.uint8)), p.spec
    kept = batch.load_tuning(path)
    assert len(kept) == len(tuned)
    tuned2, source2, _ = batch.autotune(pairs, torch, lanes=(), sample=128, log=logs.append)
    assert set(source2.values()) == {"kept"}
    assert {e: {k: v for k, v in r.items() if k != "effective"} for e, r in tuned2.items()} == \
        {e: {k: v for k, v in r.items() if k != "effective"} for e, r in tuned.items()}
    for p in pairs:
        p.close()

"""CPU checks of the C ABI library (no device calls) and of the host-side helpers."""
import os
import re

import numpy as np

from proddl2_amd import pdfft, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="pdfft.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(pd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pdfft.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(pdfft.EXPORTS)


def test_library_exports_every_proposal_symbol():
    from proddl2_amd import proposal
    lib = pdfft.load()
    names = declared_symbols("pdprop.h")
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(proposal.EXPORTS)
    assert lib.pd_prop_n_params() == proposal.LAYOUT["n"] == 167953


def test_peak_struct_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "pdfft.h")).read()
    body = re.search(r"typedef struct pd_peak \{(.*?)\} pd_peak;", text, re.S).group(1)
    fields = re.findall(r"^\s*\w+\s+(\w+)(?:\[\d\])?;", body, re.M)
    assert fields == ["rot", "rank", "s", "offset", "value", "reserved", "value64"]
    assert list(pdfft.PEAK_DTYPE.names) == fields
    assert pdfft.PEAK_DTYPE.itemsize == 40


def test_scan_flags_match_header():
    text = open(os.path.join(ROOT, "include", "pdfft.h")).read()
    flags = {k: int(v) for k, v in re.findall(r"#define PD_SCAN_([A-Z_]+) (\d+)", text)}
    assert set(flags) == {"SPLIT_PLANES", "ONE_LANE", "SPLIT_INVERSE", "SPLIT_FORWARD"}
    for name, value in flags.items():
        assert getattr(pdfft, "SCAN_" + name) == value, name
    assert len(set(flags.values())) == len(flags) and all(v & (v - 1) == 0 for v in flags.values())


def test_synthetic_globule_constraints():
    cx = synth.make_complex(1500, 450, seed=3)
    for mol, n in ((cx.receptor, 1500), (cx.ligand, 450)):
        p = mol.pos
        assert p.shape == (n, 3)
        d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
        np.fill_diagonal(d2, np.inf)
        assert d2.min() >= 9.0 - 1e-9  # min pair distance 3.0 A
        radius = (3 * n / (4 * np.pi * 0.05)) ** (1 / 3)
        centre = p.mean(0) if mol is cx.receptor else synth.diameter_centre(p)
        assert np.linalg.norm(p - centre, axis=1).max() <= radius + 3.5
        assert set(np.unique(mol.charge)) <= {-1.0, 0.0, 1.0}
        assert 0 <= mol.types.min() and mol.types.max() < 20
    assert np.allclose(synth.diameter_centre(cx.ligand.pos), 0, atol=1e-12)


def test_rotation_set_is_orthonormal_and_seeded():
    R = synth.rotation_set(200, seed=5).reshape(-1, 3, 3)
    assert np.allclose(R[0], np.eye(3))
    assert np.allclose(np.einsum("nij,nkj->nik", R, R), np.eye(3), atol=1e-13)
    assert np.allclose(np.linalg.det(R), 1)
    assert np.array_equal(R, synth.rotation_set(200, seed=5).reshape(-1, 3, 3))


def test_four_term_ligand_weights_are_independent():
    cx = synth.make_complex(400, 300, seed=1)
    _, w = synth.terms_for(cx, 4)
    assert np.linalg.matrix_rank(w) == 4


def test_errors_map_to_status_codes():
    # Argument validation happens on the host before any device call.
    lib = pdfft.load()
    d = pdfft.ScanDesc()
    d.n[:] = [32, 32, 34]
    d.grid_step, d.n_terms, d.max_n_trans, d.max_n_trans_inp = 1.0, 2, 4, 4
    import ctypes as C
    h = C.c_void_p()
    assert lib.pd_scan_create(C.byref(d), C.byref(h)) == 5  # 34 = 2 x 17: not a findBestSize edge
    assert b"factor" in lib.pd_last_error()
    d.n[:] = [51, 32, 32]  # 3 x 17
    assert lib.pd_scan_create(C.byref(d), C.byref(h)) == 5
    d.n[0] = 32
    d.max_n_trans = 40
    assert lib.pd_scan_create(C.byref(d), C.byref(h)) == 5
    d.max_n_trans = 4
    d.grid_step = -1.0
    assert lib.pd_scan_create(C.byref(d), C.byref(h)) == 1
    assert b"grid_step" in lib.pd_last_error()


def test_graph_and_handle_argument_errors_without_device():
    """Argument checks of the graph and handle entry points run before any device
    call: a null scan handle in the list is refused (not dereferenced)."""
    import ctypes as C
    lib = pdfft.load()
    P = C.c_void_p
    scans = (P * 2)(None, None)
    nrot = (C.c_int * 2)(4, 4)
    ptrs = (P * 2)(1, 1)
    lane = (C.c_int * 2)(0, 0)
    g = P()
    assert lib.pd_graph_create(2, scans, nrot, ptrs, ptrs, ptrs, lane, 1, C.byref(g)) == 1
    assert b"null scan" in lib.pd_last_error()
    assert lib.pd_graph_create(0, scans, nrot, ptrs, ptrs, ptrs, lane, 1, C.byref(g)) == 1
    assert lib.pd_scan_destroy(None) == 0
    assert lib.pd_scan_set_batch(None, 16) == 1
    assert lib.pd_graph_launch(None) == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before any
    device work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)

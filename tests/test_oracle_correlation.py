"""Pins of the oracle's projection, correlation and scan (closed forms, brute force)."""
import os

import numpy as np
import pytest

from proddl2_amd import pdfft, synth


def test_single_atom_projection_closed_form(oracle):
    # One receptor atom: every cell centre within the cutoff receives f(max(r2, minDist2)),
    # every other cell 0 (Grid/project.hpp:143-159, docking_molforce.hpp:307-310).
    n, h = (20, 22, 24), 1.0
    c = np.array([[0.37, -0.81, 1.13]])
    e4s12, as6 = 2.5, 0.7
    cut = 4.3
    g = oracle.project_receptor(n, h, c, pdfft.FIELD_LJ_REP, np.array([[e4s12, as6]]), cut)
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1)
    centres = np.stack([oracle.geometry_centered(h, k)[1] + h * idx[..., d] for d, k in enumerate(n)], -1)
    r2 = ((centres - c[0]) ** 2).sum(-1)
    md2 = (h / 4.0) ** 2
    r2c = np.maximum(r2, md2)
    exp = np.where(r2 <= cut * cut, e4s12 / (r2c ** 3 + as6) ** 2, 0.0)
    assert np.allclose(g, exp, rtol=1e-13, atol=0)
    assert (g != 0).sum() == (r2 <= cut * cut).sum()


def test_field_kinds_closed_forms(oracle):
    n, h = (16, 16, 16), 1.0
    c = np.array([[0.2, 0.1, -0.3]])
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1)
    zsc = oracle.geometry_centered(h, 16)[1]
    r2 = ((zsc + h * idx - c[0]) ** 2).sum(-1)
    cut = 5.0
    inside = r2 <= cut * cut
    cases = [
        (pdfft.FIELD_LJ_ATT, [3.0, 0.5], lambda q: -3.0 / (q ** 3 + 0.5), -1.0),
        (pdfft.FIELD_TWOSTEP, [3.0, -1.0, 9.0], lambda q: np.where(q <= 9.0, 3.0, -1.0), -1.0),
        (pdfft.FIELD_COULOMB, [83.0], lambda q: 83.0 / q, 2.0),
        (pdfft.FIELD_DESOLV, [0.6], lambda q: 0.6 * (1 - q / 25.0) ** 2, -1.0),
    ]
    for kind, p, f, md in cases:
        g = oracle.project_receptor(n, h, c, kind, np.array([p]), cut, md)
        m2 = (h / 4) ** 2 if md < 0 else md * md
        exp = np.where(inside, f(np.maximum(r2, m2)), 0.0)
        assert np.allclose(g, exp, rtol=1e-13, atol=1e-300), kind


def test_projection_is_additive_over_atoms(oracle):
    cx = synth.make_complex(120, 30, seed=2)
    n, h = (24, 24, 24), 1.5
    rt, w = synth.terms_for(cx, 4)
    pos = cx.receptor.pos
    for kind, cut, md, p in rt:
        full = oracle.project_receptor(n, h, pos, kind, p, cut, md)
        a = oracle.project_receptor(n, h, pos[:60], kind, p[:60], cut, md)
        b = oracle.project_receptor(n, h, pos[60:], kind, p[60:], cut, md)
        assert np.allclose(full, a + b, rtol=1e-12, atol=1e-12)


def test_ljcomp_reproduces_pairwise_lj(oracle):
    # docking_molforce.hpp:461-581 + potentials_inc.hpp:280-287: with alpha = 0,
    # RepR*RepL + AttrR*AttrL equals the 6-12 LJ with geometric mixing.
    rng = np.random.default_rng(0)
    eps_r, sig_r = rng.uniform(0.1, 0.5, 5), rng.uniform(2.5, 4.0, 5)
    eps_l, sig_l = rng.uniform(0.1, 0.5, 4), rng.uniform(2.5, 4.0, 4)
    rep, att, wr, wa = oracle.ljcomp_params(eps_r, sig_r, np.zeros(5), eps_l, sig_l)
    for a in range(5):
        for j in range(4):
            for r in (3.0, 3.7, 5.1):
                r2 = r * r
                e = rep[a, 0] / (r2 ** 3 + rep[a, 1]) ** 2 * wr[j] - att[a, 0] / (r2 ** 3 + att[a, 1]) * wa[j]
                eps = np.sqrt(eps_r[a] * eps_l[j])
                sig = np.sqrt(sig_r[a] * sig_l[j])
                exact = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
                assert abs(e - exact) <= 1e-12 * max(1.0, abs(exact))


def test_ljcomp_library_matches_oracle(oracle):
    rng = np.random.default_rng(5)
    args = (rng.uniform(0.1, 0.5, 40), rng.uniform(2.5, 4.0, 40), rng.uniform(0, 1, 40),
            rng.uniform(0.1, 0.5, 30), rng.uniform(2.5, 4.0, 30))
    a = oracle.ljcomp_params(*args)
    b = pdfft.ljcomp_params(*args)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _softcore_lj(r, eps, sig, alpha):
    """Soft-core LJ 4 eps (s12 / A^2 - s6 / A), A = r^6 + alpha s6, and dU/dr (analytic)."""
    s6 = sig ** 6
    A = r ** 6 + alpha * s6
    u = 4 * eps * (s6 * s6 / A ** 2 - s6 / A)
    du = 4 * eps * (-2 * s6 * s6 / A ** 3 + s6 / A ** 2) * 6 * r ** 5
    return u, du


@pytest.mark.parametrize("mix", [0, 1])
def test_ljavg_field_closed_form_with_shifted_force(oracle, mix):
    """ljAvg (MolForceLJ, docking_molforce.hpp:377-450; SoftCoreLJ with
    CUT_POLICY_SHIFT_FORCE, potentials_inc.hpp:185-219): the receptor field of
    atom a is U(r) - U(c) - U'(c) (r - c) for the soft-core LJ of atom a mixed
    with the ligand's average atom; it and its slope vanish at the cutoff c."""
    rng = np.random.default_rng(3)
    eps_r, sig_r, al_r = rng.uniform(0.1, 0.8, 3), rng.uniform(3.0, 3.9, 3), rng.uniform(0.0, 0.6, 3)
    eps_l, sig_l = rng.uniform(0.1, 0.8, 7), rng.uniform(3.0, 3.9, 7)
    cut = 8.5
    par, w = oracle.ljavg_params(eps_r, sig_r, al_r, eps_l, sig_l, mix, cut)
    assert np.all(w == 1)
    for a in range(3):
        sig = (sig_r[a] + sig_l.mean()) / 2 if mix == 0 else np.sqrt(sig_r[a] * sig_l.mean())
        eps = np.sqrt(eps_r[a] * eps_l.mean())
        uc, duc = _softcore_lj(cut, eps, sig, al_r[a])
        # single atom at a cell-centre offset: every cell within the cutoff
        n, h = (26, 26, 26), 0.7
        c = np.array([[0.11, -0.23, 0.05]])
        g = oracle.project_receptor(n, h, c, pdfft.FIELD_LJ_FULL, par[a:a + 1], cut)
        idx = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1)
        zsc = oracle.geometry_centered(h, 26)[1]
        r2 = ((zsc + h * idx - c[0]) ** 2).sum(-1)
        r = np.sqrt(np.maximum(r2, (h / 4) ** 2))
        u, _ = _softcore_lj(r, eps, sig, al_r[a])
        exp = np.where(r2 <= cut * cut, u - uc - duc * (r - cut), 0.0)
        scale = np.abs(exp).max()
        assert np.abs(g - exp).max() <= 1e-11 * scale, a
        # the shifts: energy and force vanish at the cutoff
        p0 = par[a]
        A = cut ** 6 + p0[2]
        f_c = (p0[0] / A - p0[1]) / A + p0[3] + p0[4] * cut
        assert abs(f_c) <= 1e-12 * scale
        assert abs(p0[4] + duc) <= 1e-10 * abs(duc)


def test_ljavg_library_matches_oracle(oracle):
    rng = np.random.default_rng(6)
    args = (rng.uniform(0.1, 0.5, 40), rng.uniform(2.5, 4.0, 40), rng.uniform(0, 1, 40),
            rng.uniform(0.1, 0.5, 30), rng.uniform(2.5, 4.0, 30))
    for mix in (0, 1):
        a = oracle.ljavg_params(*args, mix, 9.0)
        b = pdfft.ljavg_params(*args, mix=mix, cutoff=9.0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), mix


def _case(oracle, n, h, T, seed, n_rec=150, n_lig=30):
    cx = synth.make_complex(n_rec, n_lig, seed)
    rt, w = synth.terms_for(cx, T)
    R = np.stack([oracle.project_receptor(n, h, cx.receptor.pos, k, p, cut, md) for (k, cut, md, p) in rt])
    return cx, w, R


@pytest.mark.parametrize("n", [(16, 18, 20), (15, 14, 13), (12, 12, 12)])
def test_correlation_equals_direct_sum(oracle, n):
    # E(s) = C(-s) = sum_t sum_j w_tj R_t(l_j + s), brute force at every offset.
    h, T = 1.6, 2
    cx, w, R = _case(oracle, n, h, T, seed=4, n_rec=80, n_lig=12)
    S = oracle.receptor_spectra(R)
    rot = synth.rotation_set(2, seed=9)[1]
    q = oracle.rotate_points(rot, cx.ligand.pos)
    L = np.stack([oracle.project_ligand(n, h, q, w[t]) for t in range(T)])
    C = oracle.correlate(S, L)
    cells = oracle.cells(n, h, q)
    rng = np.random.default_rng(1)
    for x in [(0, 0, 0)] + [tuple(rng.integers(0, k) for k in n) for _ in range(25)]:
        s = [(-xi) % ni for xi, ni in zip(x, n)]
        e = oracle.direct_energy(R, cells, w, s)
        assert abs(C[x] - e) <= 1e-9 * max(1.0, abs(e))


def test_delta_ligand_returns_receptor(oracle):
    # docking_app.hpp:546-553: a single unit weight correlates back to the receptor grid.
    n, h = (10, 12, 14), 1.0
    rng = np.random.default_rng(2)
    R = rng.standard_normal((1, *n))
    L = np.zeros((1, *n))
    c0 = (3, 5, 7)
    L[(0, *c0)] = 1.0
    C = oracle.correlate(oracle.receptor_spectra(R), L)
    x = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1)
    src = tuple(((np.array(c0) - x) % np.array(n))[..., d] for d in range(3))
    assert np.allclose(C, R[0][src], atol=1e-12)


def test_doubled_term_doubles_energy_same_poses(oracle):
    # MolForceLJCompTest (docking_molforce.hpp:605-626): two copies of a term give
    # the same translations with doubled energies.
    n, h = (20, 20, 20), 1.5
    cx, w, R = _case(oracle, n, h, 2, seed=6)
    S1 = oracle.receptor_spectra(R[:1])
    S2 = oracle.receptor_spectra(np.concatenate([R[:1], R[:1]]))
    rots = synth.rotation_set(3, seed=2)
    a = oracle.scan(n, h, S1, cx.ligand.pos, w[:1], rots, K=6)
    b = oracle.scan(n, h, S2, cx.ligand.pos, np.concatenate([w[:1], w[:1]]), rots, K=6)
    assert (a[1] == b[1]).all()
    assert np.allclose(2 * a[3], b[3], rtol=1e-12)


def test_scan_equals_composed_steps(oracle):
    n, h, T, K = (18, 16, 20), 1.5, 4, 7
    cx, w, R = _case(oracle, n, h, T, seed=8)
    S = oracle.receptor_spectra(R)
    rots = synth.rotation_set(4, seed=3)
    nout, off, s, val, grids = oracle.scan(n, h, S, cx.ligand.pos, w, rots, K=K, n_threads=3, want_grids=True)
    for r in range(4):
        q = oracle.rotate_points(rots[r], cx.ligand.pos)
        L = np.stack([oracle.project_ligand(n, h, q, w[t]) for t in range(T)])
        C = oracle.correlate(S, L)
        assert np.array_equal(C, grids[r])
        o2, s2, _, v2 = oracle.select(C, h, 100 * K, 1e16, K)
        assert (o2 == off[r]).all() and (s2 == s[r]).all() and np.array_equal(v2, val[r])


def test_ligand_outside_grid_raises(oracle):
    n, h = (8, 8, 8), 1.0
    with pytest.raises(ValueError):
        oracle.project_ligand(n, h, np.array([[10.0, 0, 0]]), np.array([1.0]))


def test_scan_record_round_trip(oracle, tmp_path):
    # IO/test_dock_io_bin.cpp:41-71: two records, read back, EOF returns no more.
    p = str(tmp_path / "scan.bin")
    rot = np.eye(3)
    tran = np.array([[2.0, 2, 2], [0.0, 0, 0]])
    oracle.write_scan_record(p, rot, tran, [1.0, 2.0], append=False)
    oracle.write_scan_record(p, rot, tran, [1.0, 3.0], append=True)
    assert os.path.getsize(p) == 2 * (4 + 72 + 8 + 2 * 32)
    rots, counts, tr, vals = oracle.read_scan_records(p)
    assert len(counts) == 2 and list(counts) == [2, 2]
    assert vals[0, 1] == 2.0 and vals[1, 1] == 3.0
    assert np.array_equal(rots[0].reshape(3, 3), rot)
    raw = open(p, "rb").read()
    assert int.from_bytes(raw[:4], "little") == 2014011401


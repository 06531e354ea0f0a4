"""ctypes binding of the FP64 CPU oracle (oracle/_build/libpdoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product library never uses it.
Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpdoracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        i, d, s = C.c_int, C.c_double, C.c_char_p
        sig = {
            "pdo_field_nparams": ([i], i),
            "pdo_find_first_prime_factoring": ([i, i, ip, i, ip, ip], i),
            "pdo_find_best_size": ([i, ip], i),
            "pdo_geometry_centered": ([d, i, dp, dp], None),
            "pdo_to_logical": ([d, d, d], i),
            "pdo_grid_size_from_box": ([d, d, d], i),
            "pdo_euler_to_matrix": ([dp, dp], None),
            "pdo_matrix_to_euler": ([dp, dp, ip], None),
            "pdo_load_rotational_grid": ([s, dp, i], i),
            "pdo_rotate_points": ([dp, i, dp, dp], None),
            "pdo_matmul3": ([dp, dp, dp], None),
            "pdo_project_receptor": ([ip, d, i, dp, i, dp, d, d, dp], None),
            "pdo_project_ligand": ([ip, d, i, dp, dp, dp], i),
            "pdo_cells": ([ip, d, i, dp, ip], None),
            "pdo_ljcomp_params": ([i, dp, dp, dp, i, dp, dp, dp, dp, dp, dp], None),
            "pdo_ljavg_params": ([i, dp, dp, dp, i, dp, dp, i, d, dp, dp], None),
            "pdo_fft1": ([i, i, dp], None),
            "pdo_rfft3": ([ip, dp, dp], None),
            "pdo_irfft3": ([ip, dp, dp], None),
            "pdo_correlate": ([ip, i, dp, dp, dp], None),
            "pdo_select": ([ip, d, dp, i, d, i, ip, ip, dp, dp], i),
            "pdo_direct_energy": ([ip, i, dp, i, ip, dp, ip], d),
            "pdo_scan": ([ip, d, i, dp, i, dp, dp, i, dp, i, d, i, i, ip, ip, ip, dp, dp], i),
            "pdo_tran_symm": ([i, dp, dp, ip, dp, dp, dp, i, d, d, dp], i),
            "pdo_tran_clust": ([i, dp, dp, ip, d, i, i], i),
            "pdo_select_post": ([ip, d, dp, i, d, i, dp, dp, dp, i, d, d, d, ip, ip, dp], i),
            "pdo_write_scan_record": ([s, i, dp, i, dp, dp], i),
            "pdo_read_scan_records": ([s, i, i, dp, ip, dp, dp], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _n(n):
    return np.ascontiguousarray(n, np.int32)


def find_first_prime_factoring(start, end, primes):
    pr = np.ascontiguousarray(primes, np.int32)
    ex = np.zeros(len(pr), np.int32)
    res = np.zeros(1, np.int32)
    ok = load().pdo_find_first_prime_factoring(int(start), int(end), _i(pr), len(pr), _i(ex), _i(res))
    return bool(ok), int(res[0]), ex


def find_best_size(start):
    res = np.zeros(1, np.int32)
    ok = load().pdo_find_best_size(int(start), _i(res))
    return int(res[0]) if ok else -1


def geometry_centered(h, n):
    zh, zc = C.c_double(), C.c_double()
    load().pdo_geometry_centered(float(h), int(n), C.byref(zh), C.byref(zc))
    return zh.value, zc.value


def to_logical(h, zero_h, x):
    return load().pdo_to_logical(float(h), float(zero_h), float(x))


def grid_size_from_box(lo, hi, h):
    return load().pdo_grid_size_from_box(float(lo), float(hi), float(h))


def euler_to_matrix(ang):
    a = np.ascontiguousarray(ang, np.float64)
    m = np.empty(9)
    load().pdo_euler_to_matrix(_d(a), _d(m))
    return m.reshape(3, 3)


def matrix_to_euler(m):
    mm = np.ascontiguousarray(m, np.float64).reshape(9)
    a = np.empty(3)
    u = C.c_int()
    load().pdo_matrix_to_euler(_d(mm), _d(a), C.byref(u))
    return a, bool(u.value)


def load_rotational_grid(path, cap=1 << 20):
    lib = load()
    n = lib.pdo_load_rotational_grid(path.encode(), None, 0)
    if n < 0:
        raise IOError(f"cannot read rotation grid {path}")
    m = np.empty((n, 9))
    lib.pdo_load_rotational_grid(path.encode(), _d(m), n)
    return m


def rotate_points(m, pts):
    mm = np.ascontiguousarray(m, np.float64).reshape(9)
    p = np.ascontiguousarray(pts, np.float64)
    out = np.empty_like(p)
    load().pdo_rotate_points(_d(mm), p.shape[0], _d(p), _d(out))
    return out


def matmul3(a, b):
    aa = np.ascontiguousarray(a, np.float64).reshape(9)
    bb = np.ascontiguousarray(b, np.float64).reshape(9)
    c = np.empty(9)
    load().pdo_matmul3(_d(aa), _d(bb), _d(c))
    return c.reshape(3, 3)


def project_receptor(n, h, pos, kind, params, cutoff, min_dist=-1.0):
    nn = _n(n)
    p = np.ascontiguousarray(pos, np.float64)
    par = np.ascontiguousarray(params, np.float64)
    g = np.empty(tuple(int(x) for x in n))
    load().pdo_project_receptor(_i(nn), float(h), p.shape[0], _d(p), int(kind), _d(par),
                                float(cutoff), float(min_dist), _d(g))
    return g


def project_ligand(n, h, pos, w):
    nn = _n(n)
    p = np.ascontiguousarray(pos, np.float64)
    ww = np.ascontiguousarray(w, np.float64)
    g = np.empty(tuple(int(x) for x in n))
    bad = load().pdo_project_ligand(_i(nn), float(h), p.shape[0], _d(p), _d(ww), _d(g))
    if bad:
        raise ValueError(f"ligand atom {bad - 1} outside the grid")
    return g


def cells(n, h, pos):
    nn = _n(n)
    p = np.ascontiguousarray(pos, np.float64)
    c = np.empty((p.shape[0], 3), np.int32)
    load().pdo_cells(_i(nn), float(h), p.shape[0], _d(p), _i(c))
    return c


def ljcomp_params(eps_r, sigma_r, alpha_r, eps_l, sigma_l):
    er, sr, ar, el, sl = (np.ascontiguousarray(x, np.float64) for x in (eps_r, sigma_r, alpha_r, eps_l, sigma_l))
    rep = np.empty((len(er), 2)); att = np.empty((len(er), 2))
    wr = np.empty(len(el)); wa = np.empty(len(el))
    load().pdo_ljcomp_params(len(er), _d(er), _d(sr), _d(ar), len(el), _d(el), _d(sl), _d(rep), _d(att),
                             _d(wr), _d(wa))
    return rep, att, wr, wa


def ljavg_params(eps_r, sigma_r, alpha_r, eps_l, sigma_l, mix, cutoff):
    er, sr, ar, el, sl = (np.ascontiguousarray(x, np.float64) for x in (eps_r, sigma_r, alpha_r, eps_l, sigma_l))
    par = np.empty((len(er), 5)); w = np.empty(len(el))
    load().pdo_ljavg_params(len(er), _d(er), _d(sr), _d(ar), len(el), _d(el), _d(sl), int(mix),
                            float(cutoff), _d(par), _d(w))
    return par, w


def fft1(x, sign):
    a = np.ascontiguousarray(x, np.complex128).copy()
    load().pdo_fft1(a.shape[0], int(sign), a.ctypes.data_as(dp))
    return a


def rfft3(grid):
    g = np.ascontiguousarray(grid, np.float64)
    n = _n(g.shape)
    out = np.empty((g.shape[0], g.shape[1], g.shape[2] // 2 + 1), np.complex128)
    load().pdo_rfft3(_i(n), _d(g), out.ctypes.data_as(dp))
    return out


def irfft3(spec, n):
    s = np.ascontiguousarray(spec, np.complex128)
    nn = _n(n)
    out = np.empty(tuple(int(x) for x in n))
    load().pdo_irfft3(_i(nn), s.ctypes.data_as(dp), _d(out))
    return out


def receptor_spectra(rec_grids):
    return np.stack([rfft3(g) for g in rec_grids])


def correlate(rec_spectra, lig_grids):
    L = np.ascontiguousarray(lig_grids, np.float64)
    S = np.ascontiguousarray(rec_spectra, np.complex128)
    n = _n(L.shape[1:])
    out = np.empty(L.shape[1:])
    load().pdo_correlate(_i(n), L.shape[0], S.ctypes.data_as(dp), _d(L), _d(out))
    return out


def select(grid, h, max_inp, max_val, n_max):
    g = np.ascontiguousarray(grid, np.float64)
    n = _n(g.shape)
    off = np.empty(n_max, np.int32)
    s = np.empty((n_max, 3), np.int32)
    tran = np.empty((n_max, 3))
    val = np.empty(n_max)
    k = load().pdo_select(_i(n), float(h), _d(g), int(max_inp), float(max_val), int(n_max), _i(off),
                          _i(s), _d(tran), _d(val))
    return off[:k], s[:k], tran[:k], val[:k]


def direct_energy(rec_grids, lig_cells, w, s):
    R = np.ascontiguousarray(rec_grids, np.float64)
    c = np.ascontiguousarray(lig_cells, np.int32)
    ww = np.ascontiguousarray(w, np.float64)
    ss = np.ascontiguousarray(s, np.int32)
    n = _n(R.shape[1:])
    return load().pdo_direct_energy(_i(n), R.shape[0], _d(R), c.shape[0], _i(c), _d(ww), _i(ss))


def scan(n, h, rec_spectra, lig_pos, w, rots, K, max_inp=None, max_val=1e16, n_threads=1,
         want_grids=False):
    """Reference worker loop over rotations; returns (n_out, offsets, s, values[, grids])."""
    nn = _n(n)
    S = np.ascontiguousarray(rec_spectra, np.complex128)
    q = np.ascontiguousarray(lig_pos, np.float64)
    ww = np.ascontiguousarray(w, np.float64)
    R = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
    nr = R.shape[0]
    T = ww.shape[0]
    nout = np.zeros(nr, np.int32)
    off = np.full((nr, K), -1, np.int32)
    s = np.zeros((nr, K, 3), np.int32)
    val = np.full((nr, K), np.nan)
    grids = np.empty((nr, *[int(x) for x in n])) if want_grids else None
    bad = load().pdo_scan(_i(nn), float(h), T, S.ctypes.data_as(dp), q.shape[0], _d(q), _d(ww), nr, _d(R),
                          int(max_inp if max_inp is not None else 100 * K), float(max_val), int(K),
                          int(n_threads), _i(nout), _i(off), _i(s), _d(val),
                          _d(grids) if want_grids else None)
    if bad:
        raise ValueError(f"rotation {bad - 1}: ligand atom outside the grid")
    if want_grids:
        return nout, off, s, val, grids
    return nout, off, s, val


IDENTITY12 = np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0])


def tran_symm(tran, values, rot, frame_a=IDENTITY12, frame_b=IDENTITY12, n_symm=2, mol_radius=10.0,
              max_rmsd=8.0):
    """TranSymm on translations (n x 3) and values in list order; returns
    (kept indices, rmsd of every input)."""
    t = np.array(tran, np.float64).reshape(-1, 3).copy()
    v = np.array(values, np.float64).copy()
    tag = np.arange(len(v), dtype=np.int32)
    r = np.ascontiguousarray(rot, np.float64).reshape(9)
    fa = np.ascontiguousarray(frame_a, np.float64).reshape(12)
    fb = np.ascontiguousarray(frame_b, np.float64).reshape(12)
    rmsd = np.empty(len(v))
    k = load().pdo_tran_symm(len(v), _d(t), _d(v), _i(tag), _d(r), _d(fa), _d(fb), int(n_symm), float(mol_radius),
                             float(max_rmsd), _d(rmsd))
    return tag[:k].copy(), rmsd


def tran_clust(tran, values, cluster_radius, max_out, cap=None):
    """TranClust on translations (n x 3) and values; returns the output's input indices."""
    t = np.array(tran, np.float64).reshape(-1, 3).copy()
    v = np.array(values, np.float64).copy()
    tag = np.arange(len(v), dtype=np.int32)
    k = load().pdo_tran_clust(len(v), _d(t), _d(v), _i(tag), float(cluster_radius), int(max_out),
                              int(cap if cap is not None else len(v)))
    return tag[:k].copy()


def select_post(grid, h, max_inp, max_val, K, rot=None, frame_a=IDENTITY12, frame_b=IDENTITY12, n_symm=0,
                mol_radius=0.0, max_rmsd=8.0, cluster_radius=0.0):
    """One rotation's selection + TranSymm + TranClust + first K; returns (offsets, s, values)."""
    g = np.ascontiguousarray(grid, np.float64)
    n = _n(g.shape)
    r = np.ascontiguousarray(np.eye(3) if rot is None else rot, np.float64).reshape(9)
    fa = np.ascontiguousarray(frame_a, np.float64).reshape(12)
    fb = np.ascontiguousarray(frame_b, np.float64).reshape(12)
    off = np.empty(K, np.int32)
    s = np.empty((K, 3), np.int32)
    val = np.empty(K)
    k = load().pdo_select_post(_i(n), float(h), _d(g), int(max_inp), float(max_val), int(K), _d(r), _d(fa), _d(fb),
                               int(n_symm), float(mol_radius), float(max_rmsd), float(cluster_radius), _i(off),
                               _i(s), _d(val))
    return off[:k], s[:k], val[:k]


def write_scan_record(path, rot, tran, values, append=True):
    r = np.ascontiguousarray(rot, np.float64).reshape(9)
    t = np.ascontiguousarray(tran, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(values, np.float64).reshape(-1)
    return load().pdo_write_scan_record(path.encode(), int(append), _d(r), t.shape[0], _d(t), _d(v))


def read_scan_records(path, cap=100000, maxn=1000):
    rots = np.empty((cap, 9)); counts = np.empty(cap, np.int32)
    tran = np.empty((cap, maxn, 3)); vals = np.empty((cap, maxn))
    k = load().pdo_read_scan_records(path.encode(), cap, maxn, _d(rots), _i(counts), _d(tran), _d(vals))
    if k < 0:
        raise IOError(path)
    return rots[:k], counts[:k], tran[:k], vals[:k]

"""ctypes binding of libpdfft.so (include/pdfft.h).

The Python host mirror of the reference worker's in-process seams:
``Scan`` wraps one FFTCorrelators + CorrelationProcessor pair
(include/PRODDL/docking_fft.hpp:236-492, 767-883).  Every call goes through the
C ABI; a missing library or a failing call raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libpdfft.so")

PD_OK = 0
STATUS_NAMES = {
    1: "invalid argument",
    2: "domain error",
    3: "io error",
    4: "device error",
    5: "unsupported",
    6: "state error",
}

FIELD_LJ_REP, FIELD_LJ_ATT, FIELD_LJ_FULL, FIELD_TWOSTEP, FIELD_COULOMB, FIELD_DESOLV = range(6)
FIELD_NPARAMS = {0: 2, 1: 2, 2: 5, 3: 3, 4: 1, 5: 1}

# Exported symbols, in the order include/pdfft.h declares them.
EXPORTS = (
    "pd_scan_create", "pd_scan_project_receptor", "pd_scan_set_receptor_grids",
    "pd_scan_get_receptor_grids", "pd_scan_set_ligand", "pd_scan_run", "pd_scan_run_device",
    "pd_scan_sync", "pd_scan_rescore64", "pd_scan_correlation_grid", "pd_scan_set_timing",
    "pd_scan_kernel_stats", "pd_scan_stream", "pd_scan_destroy", "pd_last_error",
    "pd_find_best_size", "pd_ljcomp_params", "pd_scan_set_batch", "pd_graph_create", "pd_graph_launch",
    "pd_graph_sync", "pd_graph_stream", "pd_graph_destroy", "pd_ljavg_params",
    "pd_scan_rescore64_noalloc", "pd_scan_set_postfilter", "pd_scan_candidates", "pd_scan_ligand_box",
    "pd_min_volume_box", "pd_diameter_pair", "pd_scan_set_tuning", "pd_scan_get_tuning",
)

KERNEL_NAMES = ("k_bin", "k_lig_plane_fwd", "k_col", "k_plane_inv_topk", "k_merge")


class PdError(RuntimeError):
    """A libpdfft call failed; ``status`` is the pd_status code."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class ScanDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("grid_step", C.c_double),
        ("n_terms", C.c_int),
        ("max_n_trans", C.c_int),
        ("max_n_trans_inp", C.c_int),
        ("max_val_corr", C.c_double),
        ("batch", C.c_int),
        ("device", C.c_int),
        ("flags", C.c_int),
    ]


SCAN_SPLIT_PLANES = 1
SCAN_ONE_LANE = 2
SCAN_SPLIT_INVERSE = 4
SCAN_SPLIT_FORWARD = 8


class PostFilter(C.Structure):
    _fields_ = [
        ("n_multimer", C.c_int),
        ("max_rmsd_symm", C.c_double),
        ("ligand_radius", C.c_double),
        ("frame_rec_inv", C.c_double * 12),
        ("frame_lig", C.c_double * 12),
        ("cluster", C.c_int),
        ("cluster_radius", C.c_double),
    ]


IDENTITY12 = (1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)


class Tuning(C.Structure):
    """pd_tuning: launch choices for an autotuner (0 keeps the library's rule)."""
    _fields_ = [("col_groups", C.c_int), ("col_mz_extra", C.c_int), ("lig_pass", C.c_int), ("lanes", C.c_int)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Term(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("cutoff", C.c_double),
        ("min_dist", C.c_double),
        ("params", C.POINTER(C.c_double)),
    ]


PEAK_DTYPE = np.dtype([
    ("rot", "<i4"), ("rank", "<i4"), ("s", "<i4", (3,)), ("offset", "<i4"),
    ("value", "<f4"), ("reserved", "<f4"), ("value64", "<f8"),
])
assert PEAK_DTYPE.itemsize == 40

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libpdfft.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m proddl2_amd.build` "
            "(the scan has no CPU fallback)")
    # PyTorch wheels carry their own libamdhip64; if torch is present it must be
    # loaded first so that this library binds to the same HIP runtime (two
    # runtimes in one process leave torch without devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    P = C.c_void_p
    i, d = C.c_int, C.c_double
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lib.pd_scan_create.argtypes = [C.POINTER(ScanDesc), C.POINTER(P)]
    lib.pd_scan_project_receptor.argtypes = [P, i, dp, C.POINTER(Term)]
    lib.pd_scan_set_receptor_grids.argtypes = [P, dp]
    lib.pd_scan_get_receptor_grids.argtypes = [P, dp]
    lib.pd_scan_set_ligand.argtypes = [P, i, dp, dp]
    lib.pd_scan_run.argtypes = [P, i, dp, i, P, ip]
    lib.pd_scan_run_device.argtypes = [P, i, P, i, P, P, i]
    lib.pd_scan_sync.argtypes = [P]
    lib.pd_scan_rescore64.argtypes = [P, i, dp, ip, dp]
    lib.pd_scan_rescore64_noalloc.argtypes = [P, i, dp, ip, dp]
    lib.pd_scan_correlation_grid.argtypes = [P, dp, C.POINTER(C.c_float)]
    lib.pd_scan_set_timing.argtypes = [P, i]
    lib.pd_scan_kernel_stats.argtypes = [P, i, dp, C.POINTER(C.c_longlong)]
    lib.pd_scan_stream.argtypes = [P]
    lib.pd_scan_stream.restype = P
    lib.pd_scan_ligand_box.argtypes = [P, ip, ip, ip]
    lib.pd_scan_destroy.argtypes = [P]
    lib.pd_last_error.restype = C.c_char_p
    lib.pd_find_best_size.argtypes = [i]
    lib.pd_find_best_size.restype = i
    lib.pd_ljcomp_params.argtypes = [i, dp, dp, dp, i, dp, dp, dp, dp, dp, dp]
    lib.pd_ljavg_params.argtypes = [i, dp, dp, dp, i, dp, dp, i, d, dp, dp]
    lib.pd_scan_set_batch.argtypes = [P, i]
    lib.pd_scan_set_postfilter.argtypes = [P, C.POINTER(PostFilter)]
    lib.pd_scan_candidates.argtypes = [P, i, dp, i, P, ip]
    PP = C.POINTER(P)
    lib.pd_graph_create.argtypes = [i, PP, ip, PP, PP, PP, ip, i, PP]
    lib.pd_graph_launch.argtypes = [P]
    lib.pd_graph_sync.argtypes = [P]
    lib.pd_graph_stream.argtypes = [P]
    lib.pd_graph_stream.restype = P
    lib.pd_graph_destroy.argtypes = [P]
    lib.pd_graph_destroy.restype = None
    for name in EXPORTS:
        if name not in ("pd_last_error", "pd_find_best_size", "pd_scan_stream",
                        "pd_graph_stream", "pd_graph_destroy"):
            getattr(lib, name).restype = i
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != PD_OK:
        raise PdError(status, _lib.pd_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def find_best_size(start: int) -> int:
    return load().pd_find_best_size(int(start))


def ljcomp_params(eps_r, sigma_r, alpha_r, eps_l, sigma_l):
    """Per-atom ljComp parameters (docking_molforce.hpp:461-581) via the C ABI."""
    lib = load()
    er, sr, ar = (np.ascontiguousarray(x, np.float64) for x in (eps_r, sigma_r, alpha_r))
    el, sl = (np.ascontiguousarray(x, np.float64) for x in (eps_l, sigma_l))
    rep = np.empty((len(er), 2))
    att = np.empty((len(er), 2))
    wr = np.empty(len(el))
    wa = np.empty(len(el))
    _check(lib.pd_ljcomp_params(len(er), _dp(er), _dp(sr), _dp(ar), len(el), _dp(el), _dp(sl),
                                _dp(rep), _dp(att), _dp(wr), _dp(wa)))
    return rep, att, wr, wa


def ljavg_params(eps_r, sigma_r, alpha_r, eps_l, sigma_l, mix=1, cutoff=9.0):
    """Per-atom ljAvg parameters (MolForceLJ, docking_molforce.hpp:377-450) via the C ABI:
    (params[n_rec, 5] for FIELD_LJ_FULL, ligand weights[n_lig])."""
    lib = load()
    er, sr, ar = (np.ascontiguousarray(x, np.float64) for x in (eps_r, sigma_r, alpha_r))
    el, sl = (np.ascontiguousarray(x, np.float64) for x in (eps_l, sigma_l))
    par = np.empty((len(er), 5))
    w = np.empty(len(el))
    _check(lib.pd_ljavg_params(len(er), _dp(er), _dp(sr), _dp(ar), len(el), _dp(el), _dp(sl), int(mix),
                               float(cutoff), _dp(par), _dp(w)))
    return par, w


def min_volume_box(pos):
    """Approximate minimum-volume box of the points (host code of the library;
    Geom::Bounding::Box, Geom/bounding.hpp:135-183): (axes[3, 3] with the unit
    box directions in rows, a proper rotation; lo[3]; hi[3])."""
    lib = load()
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    axes = np.empty((3, 3))
    lo = np.empty(3)
    hi = np.empty(3)
    _check(lib.pd_min_volume_box(len(p), _dp(p), _dp(axes), _dp(lo), _dp(hi)))
    return axes, lo, hi


def diameter_pair(pos):
    """Indices (i, j), i < j, of the two points farthest apart (exact)."""
    lib = load()
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    i = C.c_int()
    j = C.c_int()
    _check(lib.pd_diameter_pair(len(p), _dp(p), C.byref(i), C.byref(j)))
    return i.value, j.value


class Scan:
    """One scan handle: receptor spectra resident on the GPU, rotations streamed."""

    def __init__(self, n, grid_step: float, n_terms: int, max_n_trans: int = 10,
                 max_n_trans_inp: int | None = None, max_val_corr: float = 0.0,
                 batch: int = 64, device: int = 0, flags: int = 0):
        lib = load()
        d = ScanDesc()
        d.n[:] = [int(x) for x in n]
        d.grid_step = float(grid_step)
        d.n_terms = int(n_terms)
        d.max_n_trans = int(max_n_trans)
        d.max_n_trans_inp = int(max_n_trans_inp if max_n_trans_inp is not None else 100 * max_n_trans)
        d.max_val_corr = float(max_val_corr)
        d.batch = int(batch)
        d.device = int(device)
        d.flags = int(flags)
        h = C.c_void_p()
        _check(lib.pd_scan_create(C.byref(d), C.byref(h)))
        self._h = h
        self.n = tuple(int(x) for x in n)
        self.n_terms = int(n_terms)
        self.K = int(max_n_trans)
        self.batch = int(batch) or 64
        self.k_inp = int(d.max_n_trans_inp)
        self.grid_step = float(grid_step)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _check(_lib.pd_scan_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- receptor -----------------------------------------------------------
    def project_receptor(self, pos, terms) -> None:
        """terms: list of (kind, cutoff, min_dist, params[n_atoms, nparams])."""
        pos = np.ascontiguousarray(pos, np.float64)
        n = pos.shape[0]
        keep = []
        arr = (Term * len(terms))()
        for t, (kind, cutoff, min_dist, params) in enumerate(terms):
            p = np.ascontiguousarray(params, np.float64).reshape(n, FIELD_NPARAMS[kind])
            keep.append(p)
            arr[t].kind = int(kind)
            arr[t].cutoff = float(cutoff)
            arr[t].min_dist = float(min_dist)
            arr[t].params = _dp(p)
        _check(_lib.pd_scan_project_receptor(self._h, n, _dp(pos), arr))

    def set_receptor_grids(self, grids) -> None:
        g = np.ascontiguousarray(grids, np.float64)
        assert g.size == self.n_terms * int(np.prod(self.n))
        _check(_lib.pd_scan_set_receptor_grids(self._h, _dp(g)))

    def receptor_grids(self) -> np.ndarray:
        out = np.empty((self.n_terms, *self.n))
        _check(_lib.pd_scan_get_receptor_grids(self._h, _dp(out)))
        return out

    # -- ligand -------------------------------------------------------------
    def set_ligand(self, pos, weights) -> None:
        pos = np.ascontiguousarray(pos, np.float64)
        w = np.ascontiguousarray(weights, np.float64).reshape(self.n_terms, pos.shape[0])
        _check(_lib.pd_scan_set_ligand(self._h, pos.shape[0], _dp(pos), _dp(w)))

    # -- scan ---------------------------------------------------------------
    def run(self, rots, rot_index0: int = 0):
        """Returns (peaks[n_rot, K] structured, n_out[n_rot])."""
        r = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
        nr = r.shape[0]
        peaks = np.zeros((nr, self.K), PEAK_DTYPE)
        nout = np.zeros(nr, np.int32)
        _check(_lib.pd_scan_run(self._h, nr, _dp(r), int(rot_index0),
                                peaks.ctypes.data_as(C.c_void_p), _ip(nout)))
        return peaks, nout

    def run_device(self, n_rot: int, d_rots: int, d_peaks: int, d_nout: int, rot_index0: int = 0,
                   sync: bool = False) -> None:
        """Device pointers (ints); asynchronous on the handle's stream unless sync."""
        _check(_lib.pd_scan_run_device(self._h, int(n_rot), C.c_void_p(d_rots), int(rot_index0),
                                       C.c_void_p(d_peaks), C.c_void_p(d_nout), int(bool(sync))))

    def sync(self) -> None:
        _check(_lib.pd_scan_sync(self._h))

    def rescore64(self, rots, s) -> np.ndarray:
        r = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
        sv = np.ascontiguousarray(s, np.int32).reshape(-1, 3)
        out = np.empty(r.shape[0])
        _check(_lib.pd_scan_rescore64(self._h, r.shape[0], _dp(r), _ip(sv), _dp(out)))
        return out

    def correlation_grid(self, rot) -> np.ndarray:
        r = np.ascontiguousarray(rot, np.float64).reshape(9)
        out = np.empty(self.n, np.float32)
        _check(_lib.pd_scan_correlation_grid(self._h, _dp(r), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    # -- timing -------------------------------------------------------------
    def set_timing(self, enable: bool) -> None:
        _check(_lib.pd_scan_set_timing(self._h, int(bool(enable))))

    def kernel_stats(self):
        res = {}
        for k, name in enumerate(KERNEL_NAMES):
            ms = C.c_double()
            n = C.c_longlong()
            _check(_lib.pd_scan_kernel_stats(self._h, k, C.byref(ms), C.byref(n)))
            res[name] = (ms.value, n.value)
        return res

    def ligand_box(self):
        """(lo[3], p[3], stored x planes per rotation and term)."""
        lo = np.zeros(3, np.int32)
        p = np.zeros(3, np.int32)
        planes = np.zeros(1, np.int32)
        _check(_lib.pd_scan_ligand_box(self._h, _ip(lo), _ip(p), _ip(planes)))
        return lo, p, int(planes[0])

    def stream(self) -> int:
        return _lib.pd_scan_stream(self._h) or 0

    def set_batch(self, batch: int) -> None:
        _check(_lib.pd_scan_set_batch(self._h, int(batch)))
        self.batch = int(batch)

    def set_tuning(self, col_groups=0, col_mz_extra=0, lig_pass=0, lanes=0) -> None:
        """Launch choices (pd_scan_set_tuning); all 0: the library's rules."""
        t = Tuning(int(col_groups), int(col_mz_extra), int(lig_pass), int(lanes))
        _check(_lib.pd_scan_set_tuning(self._h, C.byref(t)))

    def tuning(self, batch=None) -> dict:
        """The choices in effect for a launch of `batch` rotations (default: the handle's batch)."""
        t = Tuning()
        _check(_lib.pd_scan_get_tuning(self._h, int(batch or self.batch), C.byref(t)))
        return t.as_dict()

    def set_postfilter(self, n_multimer=0, max_rmsd_symm=8.0, ligand_radius=0.0, frame_rec_inv=IDENTITY12,
                       frame_lig=IDENTITY12, cluster=False, cluster_radius=5.0) -> None:
        """TranSymm (n_multimer > 1) and / or TranClust (cluster) after the selection;
        both off clears them."""
        pf = PostFilter()
        pf.n_multimer = int(n_multimer)
        pf.max_rmsd_symm = float(max_rmsd_symm)
        pf.ligand_radius = float(ligand_radius)
        pf.frame_rec_inv[:] = [float(x) for x in np.asarray(frame_rec_inv, float).reshape(12)]
        pf.frame_lig[:] = [float(x) for x in np.asarray(frame_lig, float).reshape(12)]
        pf.cluster = int(bool(cluster))
        pf.cluster_radius = float(cluster_radius)
        _check(_lib.pd_scan_set_postfilter(self._h, C.byref(pf)))

    def candidates(self, rots, rot_index0: int = 0):
        """The max_n_trans_inp best translations of each rotation, before post-filters:
        (peaks[n_rot, k_inp] structured, n_out[n_rot])."""
        r = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
        nr = r.shape[0]
        peaks = np.zeros((nr, self.k_inp), PEAK_DTYPE)
        nout = np.zeros(nr, np.int32)
        _check(_lib.pd_scan_candidates(self._h, nr, _dp(r), int(rot_index0), peaks.ctypes.data_as(C.c_void_p),
                                       _ip(nout)))
        return peaks, nout

    @property
    def handle(self) -> int:
        return self._h.value


class Graph:
    """Several scans in one HIP graph (pd_graph_*), built from kernel nodes.

    pairs: list of (Scan, n_rot, d_rots, d_peaks, d_n_out, lane) with device
    pointers as ints; the buffers must stay alive as long as the graph.  The
    graph keeps its scans alive: a scan cannot be closed while a graph holds it."""

    def __init__(self, pairs, n_lanes: int):
        lib = load()
        n = len(pairs)
        P = C.c_void_p
        scans = (P * n)(*[p[0].handle for p in pairs])
        nrot = (C.c_int * n)(*[int(p[1]) for p in pairs])
        rots = (P * n)(*[int(p[2]) for p in pairs])
        peaks = (P * n)(*[int(p[3]) for p in pairs])
        nout = (P * n)(*[int(p[4]) for p in pairs])
        lane = (C.c_int * n)(*[int(p[5]) for p in pairs])
        g = P()
        _check(lib.pd_graph_create(n, scans, nrot, rots, peaks, nout, lane, int(n_lanes), C.byref(g)))
        self._g = g
        self._scans = [p[0] for p in pairs]

    def launch(self) -> None:
        _check(_lib.pd_graph_launch(self._g))

    def sync(self) -> None:
        _check(_lib.pd_graph_sync(self._g))

    def close(self) -> None:
        if getattr(self, "_g", None) is not None and self._g.value:
            _lib.pd_graph_destroy(self._g)
            self._g = None
            self._scans = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Learned rotation proposal: ctypes binding of include/pdprop.h, training and
pruned scanning.

The reference worker correlates every rotation of its grid
(include/PRODDL/docking_app.hpp:487-513).  Here a small network, trained on
the best exhaustive energy of each rotation of many synthetic pairs, ranks the
rotations of a new complex; only the top fraction is correlated
(pd_scan_run_device on the gathered rotations), and the exhaustive scan is the
fallback and the measure of recall (the share of its top poses kept).

Everything runs through libpdfft.so; there is no CPU path.

    python -m proddl2_amd.proposal train --out proddl2_amd/data/proposal_weights.npy
    python -m proddl2_amd.proposal eval --weights proddl2_amd/data/proposal_weights.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import pdfft, synth

_PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(_PKG, "data", "proposal_weights.npy")
EXPORTS = (
    "pd_prop_create", "pd_prop_destroy", "pd_prop_set_stream", "pd_prop_stream", "pd_prop_sync",
    "pd_prop_set_complex", "pd_prop_get_maps", "pd_prop_predict_device", "pd_prop_train_step_device",
    "pd_prop_gradients", "pd_prop_n_params", "pd_prop_get_params", "pd_prop_set_params",
    "pd_prop_select_device", "pd_prop_gather_rotations", "pd_prop_remap_peaks",
)

# Training set and schedule (synthetic pairs with the small-complex sizes,
# scored on the production grid so that training and use share one geometry).
TRAIN = dict(n_pairs=200, n_rec=1800, n_lig=450, n_rot=3000, seed0=1000, rot_seed0=5000,
             grid=(128, 128, 128), h=1.2, n_terms=4, epochs=12, lr=1e-3)


def _layout():
    """Offsets of the flat parameter vector (include/pdprop.h, pd_prop_get_params)."""
    c, cin, hid, nf = 8, 2, 384, 24
    stack = c * cin * 27 + c + 2 * (c * c * 27 + c)
    w1 = 2 * stack
    b1 = w1 + hid * (nf + 6)
    w2 = b1 + hid
    b2 = w2 + hid * hid
    w3 = b2 + hid
    return dict(conv_stack=stack, w1=w1, b1=b1, w2=w2, b2=b2, w3=w3, b3=w3 + hid, n=w3 + hid + 1,
                hid=hid, nin=nf + 6, nf=nf)


LAYOUT = _layout()


class PropDesc(C.Structure):
    _fields_ = [("max_rot", C.c_int), ("device", C.c_int), ("seed", C.c_uint)]


_bound = False


def _lib():
    global _bound
    lib = pdfft.load()
    if _bound:
        return lib
    P, i, f = C.c_void_p, C.c_int, C.c_float
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    lib.pd_prop_create.argtypes = [C.POINTER(PropDesc), C.POINTER(P)]
    lib.pd_prop_destroy.argtypes = [P]
    lib.pd_prop_destroy.restype = None
    lib.pd_prop_set_stream.argtypes = [P, P]
    lib.pd_prop_stream.argtypes = [P]
    lib.pd_prop_stream.restype = P
    lib.pd_prop_sync.argtypes = [P]
    lib.pd_prop_set_complex.argtypes = [P, i, dp, i, dp]
    lib.pd_prop_get_maps.argtypes = [P, fp, fp]
    lib.pd_prop_predict_device.argtypes = [P, i, P, P]
    lib.pd_prop_train_step_device.argtypes = [P, i, P, P, f, fp]
    lib.pd_prop_gradients.argtypes = [P, i, dp, fp, fp, fp]
    lib.pd_prop_n_params.argtypes = []
    lib.pd_prop_get_params.argtypes = [P, fp]
    lib.pd_prop_set_params.argtypes = [P, fp]
    lib.pd_prop_select_device.argtypes = [P, i, P, i, P]
    lib.pd_prop_gather_rotations.argtypes = [P, i, P, P, P]
    lib.pd_prop_remap_peaks.argtypes = [P, i, i, P, P, P]
    for name in EXPORTS:
        if name not in ("pd_prop_destroy", "pd_prop_stream"):
            getattr(lib, name).restype = i
    _bound = True
    return lib


def _check(status: int) -> None:
    if status != pdfft.PD_OK:
        raise pdfft.PdError(status, pdfft.load().pd_last_error().decode())


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Proposal:
    """One network with activation buffers for up to max_rot rotations."""

    def __init__(self, max_rot: int = 30000, device: int = 0, seed: int = 0):
        lib = _lib()
        d = PropDesc(int(max_rot), int(device), int(seed))
        h = C.c_void_p()
        _check(lib.pd_prop_create(C.byref(d), C.byref(h)))
        self._h = h
        self.max_rot = int(max_rot)
        self.n_params = lib.pd_prop_n_params()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().pd_prop_destroy(self._h)
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

    # -- set-up -------------------------------------------------------------
    def set_stream(self, stream: int | None) -> None:
        _check(_lib().pd_prop_set_stream(self._h, C.c_void_p(stream or None)))

    def stream(self) -> int:
        return _lib().pd_prop_stream(self._h) or 0

    def sync(self) -> None:
        _check(_lib().pd_prop_sync(self._h))

    def set_complex(self, rec_pos, lig_pos) -> None:
        r = np.ascontiguousarray(rec_pos, np.float64)
        l = np.ascontiguousarray(lig_pos, np.float64)
        _check(_lib().pd_prop_set_complex(self._h, r.shape[0], _dp(r), l.shape[0], _dp(l)))

    def maps(self):
        rec = np.empty((2, 16, 16, 16), np.float32)
        lig = np.empty((2, 16, 16, 16), np.float32)
        _check(_lib().pd_prop_get_maps(self._h, _fp(rec), _fp(lig)))
        return rec, lig

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _check(_lib().pd_prop_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        _check(_lib().pd_prop_set_params(self._h, _fp(p)))

    def load(self, path: str = WEIGHTS) -> None:
        self.set_params(np.load(path, allow_pickle=False))

    # -- compute ------------------------------------------------------------
    def predict_device(self, n: int, d_rots: int, d_scores: int) -> None:
        _check(_lib().pd_prop_predict_device(self._h, int(n), C.c_void_p(d_rots), C.c_void_p(d_scores)))

    def train_step_device(self, n: int, d_rots: int, d_energy: int, lr: float, want_loss: bool = True):
        loss = C.c_float(float("nan"))
        _check(_lib().pd_prop_train_step_device(self._h, int(n), C.c_void_p(d_rots), C.c_void_p(d_energy),
                                                float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def gradients(self, rots, energy):
        r = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
        e = np.ascontiguousarray(energy, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        _check(_lib().pd_prop_gradients(self._h, r.shape[0], _dp(r), _fp(e), _fp(g), C.byref(loss)))
        return g, loss.value

    def select_device(self, n: int, d_scores: int, n_keep: int, d_idx: int) -> None:
        _check(_lib().pd_prop_select_device(self._h, int(n), C.c_void_p(d_scores), int(n_keep), C.c_void_p(d_idx)))

    def gather_device(self, n: int, d_idx: int, d_in: int, d_out: int) -> None:
        _check(_lib().pd_prop_gather_rotations(self._h, int(n), C.c_void_p(d_idx), C.c_void_p(d_in),
                                               C.c_void_p(d_out)))

    def remap_peaks_device(self, n: int, K: int, d_peaks: int, d_nout: int, d_idx: int) -> None:
        _check(_lib().pd_prop_remap_peaks(self._h, int(n), int(K), C.c_void_p(d_peaks), C.c_void_p(d_nout),
                                          C.c_void_p(d_idx)))


# ---------------------------------------------------------------------------
# Driving code (device buffers are torch tensors: plumbing only)
# ---------------------------------------------------------------------------
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the proposal network runs on the GPU only")
    return torch


def exhaustive_best(scan: "pdfft.Scan", rots: np.ndarray):
    """Exhaustive scan of rots on the device: (peaks[n, K], best energy per rotation)."""
    torch = _torch()
    dev = torch.device("cuda:0")
    n = rots.shape[0]
    d_rots = torch.from_numpy(np.ascontiguousarray(rots, np.float64)).to(dev)
    d_peaks = torch.zeros((n, scan.K * pdfft.PEAK_DTYPE.itemsize // 4), dtype=torch.int32, device=dev)
    d_nout = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    scan.run_device(n, d_rots.data_ptr(), d_peaks.data_ptr(), d_nout.data_ptr(), 0, sync=True)
    peaks = d_peaks.cpu().numpy().view(pdfft.PEAK_DTYPE).reshape(n, scan.K)
    return peaks, peaks["value"][:, 0].astype(np.float32)


def training_set(n_pairs=TRAIN["n_pairs"], n_rot=TRAIN["n_rot"], log=print):
    """Synthetic pairs and the exhaustive best energy of each of their rotations."""
    T = TRAIN["n_terms"]
    data = []
    t0 = time.perf_counter()
    for i in range(n_pairs):
        cx = synth.make_complex(TRAIN["n_rec"], TRAIN["n_lig"], TRAIN["seed0"] + i)
        rt, w = synth.terms_for(cx, T)
        rots = synth.rotation_set(n_rot, TRAIN["rot_seed0"] + i)
        with pdfft.Scan(TRAIN["grid"], TRAIN["h"], T, max_n_trans=1, max_val_corr=1e16) as sc:
            sc.project_receptor(cx.receptor.pos, rt)
            sc.set_ligand(cx.ligand.pos, w)
            _, best = exhaustive_best(sc, rots)
        data.append((cx.receptor.pos, cx.ligand.pos, rots, best))
        if (i + 1) % 50 == 0:
            log(f"training set: {i + 1}/{n_pairs} pairs scanned ({time.perf_counter() - t0:.1f} s)")
    return data


def train(prop: Proposal, data, epochs=TRAIN["epochs"], lr=TRAIN["lr"], seed=0, log=print):
    """Adam on one pair per step (all its rotations), pairs in seeded random order."""
    torch = _torch()
    dev = torch.device("cuda:0")
    dev_data = [(torch.from_numpy(np.ascontiguousarray(r, np.float64)).to(dev),
                 torch.from_numpy(np.ascontiguousarray(e, np.float32)).to(dev)) for _, _, r, e in data]
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(data)):
            rec, lig, rots, _ = data[i]
            prop.set_complex(rec, lig)
            d_r, d_e = dev_data[i]
            tot += prop.train_step_device(rots.shape[0], d_r.data_ptr(), d_e.data_ptr(), lr)
        history.append(tot / len(data))
        log(f"epoch {ep}: mean loss {history[-1]:.4f}")
    return history


def top_pose_rotations(peaks: np.ndarray, n_top: int = 100) -> np.ndarray:
    """Rotation index of each of the n_top lowest-energy poses (rank order, stable)."""
    flat = peaks["value"].reshape(-1)
    order = np.argsort(flat, kind="stable")[:n_top]
    return peaks["rot"].reshape(-1)[order]


def pruned_buffers(scan: "pdfft.Scan", prop: Proposal, d_rots, n_keep: int):
    """Proposal -> selection -> full scan of the kept rotations, all on the scan's
    stream.  Returns device tensors (peaks with global rotation indices, n_out, idx)."""
    torch = _torch()
    dev = d_rots.device
    n = d_rots.shape[0]
    d_scores = torch.empty(n, dtype=torch.float32, device=dev)
    d_idx = torch.empty(n_keep, dtype=torch.int32, device=dev)
    d_kept = torch.empty((n_keep, 9), dtype=torch.float64, device=dev)
    d_peaks = torch.zeros((n_keep, scan.K * pdfft.PEAK_DTYPE.itemsize // 4), dtype=torch.int32, device=dev)
    d_nout = torch.zeros(n_keep, dtype=torch.int32, device=dev)
    return d_scores, d_idx, d_kept, d_peaks, d_nout


def run_pruned(scan, prop, d_rots, bufs, n_keep: int) -> None:
    d_scores, d_idx, d_kept, d_peaks, d_nout = bufs
    n = d_rots.shape[0]
    prop.predict_device(n, d_rots.data_ptr(), d_scores.data_ptr())
    prop.select_device(n, d_scores.data_ptr(), n_keep, d_idx.data_ptr())
    prop.gather_device(n_keep, d_idx.data_ptr(), d_rots.data_ptr(), d_kept.data_ptr())
    scan.run_device(n_keep, d_kept.data_ptr(), d_peaks.data_ptr(), d_nout.data_ptr(), 0, sync=False)
    prop.remap_peaks_device(n_keep, scan.K, d_peaks.data_ptr(), d_nout.data_ptr(), d_idx.data_ptr())


def fixed_score_params() -> np.ndarray:
    """Parameters under which the network scores a rotation by minus the sum of
    its 24 standardised rotation functions: identity convolutions (last layer
    0), no learned weighting.  This is the ranking the learned part improves on."""
    p = np.zeros(LAYOUT["n"], np.float32)
    hid, nin, nf = LAYOUT["hid"], LAYOUT["nin"], LAYOUT["nf"]
    W1 = np.zeros((hid, nin), np.float32)
    W1[0, :nf] = 1.0
    W1[1, :nf] = -1.0
    W2 = np.zeros((hid, hid), np.float32)
    W2[0, 0], W2[0, 1] = 1.0, -1.0   # relu(s)
    W2[1, 0], W2[1, 1] = -1.0, 1.0   # relu(-s)
    W3 = np.zeros(hid, np.float32)
    W3[0], W3[1] = -1.0, 1.0         # -relu(s) + relu(-s) = -s
    p[LAYOUT["w1"]:LAYOUT["w1"] + W1.size] = W1.reshape(-1)
    p[LAYOUT["w2"]:LAYOUT["w2"] + W2.size] = W2.reshape(-1)
    p[LAYOUT["w3"]:LAYOUT["w3"] + W3.size] = W3
    return p


def pruned_eval(sc, prop: Proposal, d_rots, ex_peaks, t_exhaustive_s: float, keep_fracs=(0.10, 0.15),
                n_top=100, repeats=3, log=print):
    """Pruned scans of the rotations d_rots (maps of the complex already set, prop on
    the scan's stream) against the exhaustive peaks ex_peaks of the same rotations."""
    torch = _torch()
    n = d_rots.shape[0]
    top = top_pose_rotations(ex_peaks, n_top)
    top_flat = np.argsort(ex_peaks["value"].reshape(-1), kind="stable")[:n_top]
    ex_pose = set(zip(ex_peaks["rot"].reshape(-1)[top_flat].tolist(),
                      ex_peaks["offset"].reshape(-1)[top_flat].tolist()))
    res = {"n_top": n_top, "exhaustive_s": t_exhaustive_s}
    d_scores = torch.empty(n, dtype=torch.float32, device=d_rots.device)
    t_pr = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        prop.predict_device(n, d_rots.data_ptr(), d_scores.data_ptr())
        sc.sync()
        t_pr.append(time.perf_counter() - t0)
    res["proposal_s"] = float(np.median(t_pr))
    res["keep"] = []
    for frac in keep_fracs:
        n_keep = int(round(frac * n))
        bufs = pruned_buffers(sc, prop, d_rots, n_keep)
        t_all = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run_pruned(sc, prop, d_rots, bufs, n_keep)
            sc.sync()
            t_all.append(time.perf_counter() - t0)
        kept_peaks = bufs[3].cpu().numpy().view(pdfft.PEAK_DTYPE).reshape(n_keep, sc.K)
        kept = np.zeros(n, bool)
        kept[bufs[1].cpu().numpy()] = True
        # the pruned search's own top poses against the exhaustive ones (rotation, cell)
        pr_flat = np.argsort(kept_peaks["value"].reshape(-1), kind="stable")[:n_top]
        pr_pose = set(zip(kept_peaks["rot"].reshape(-1)[pr_flat].tolist(),
                          kept_peaks["offset"].reshape(-1)[pr_flat].tolist()))
        recall = len(pr_pose & ex_pose) / n_top
        t = float(np.median(t_all))
        res["keep"].append({"fraction": frac, "n_keep": n_keep, "recall": recall,
                            "recall_by_rotation": float(kept[top].mean()), "pruned_s": t,
                            "speedup": t_exhaustive_s / t})
        log(f"keep {frac:.2f}: recall@{n_top} {recall:.2f}, pruned {t * 1e3:.1f} ms vs exhaustive "
            f"{t_exhaustive_s * 1e3:.1f} ms -> {t_exhaustive_s / t:.2f}x")
    return res


def evaluate(prop: Proposal, workload="exhaustive-128", seed=2024, keep_fracs=(0.10, 0.15), n_top=100,
             repeats=3, log=print):
    """Exhaustive vs pruned scan of one complex: recall of the exhaustive top poses and times."""
    torch = _torch()
    dev = torch.device("cuda:0")
    wl = dict(synth.WORKLOADS[workload])
    cx = synth.make_complex(wl["n_rec"], wl["n_lig"], seed)
    rt, w = synth.terms_for(cx, wl["n_terms"])
    rots = synth.rotation_set(wl["n_rot"], seed + 1)
    n = rots.shape[0]
    with pdfft.Scan(wl["n"], wl["h"], wl["n_terms"], max_n_trans=wl["K"], max_val_corr=1e16) as sc:
        sc.project_receptor(cx.receptor.pos, rt)
        sc.set_ligand(cx.ligand.pos, w)
        prop.set_stream(sc.stream())
        try:
            t0 = time.perf_counter()
            prop.set_complex(cx.receptor.pos, cx.ligand.pos)
            maps_ms = (time.perf_counter() - t0) * 1e3
            d_rots = torch.from_numpy(rots).to(dev)
            d_peaks = torch.zeros((n, sc.K * pdfft.PEAK_DTYPE.itemsize // 4), dtype=torch.int32, device=dev)
            d_nout = torch.zeros(n, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            t_ex = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                sc.run_device(n, d_rots.data_ptr(), d_peaks.data_ptr(), d_nout.data_ptr(), 0, sync=True)
                t_ex.append(time.perf_counter() - t0)
            ex_peaks = d_peaks.cpu().numpy().view(pdfft.PEAK_DTYPE).reshape(n, sc.K)
            res = pruned_eval(sc, prop, d_rots, ex_peaks, float(np.median(t_ex)), keep_fracs, n_top, repeats, log)
        finally:
            prop.set_stream(None)
    res.update({"workload": workload, "seed": seed, "rotations": n, "maps_ms": maps_ms})
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=TRAIN["n_pairs"])
    t.add_argument("--rotations", type=int, default=TRAIN["n_rot"])
    t.add_argument("--epochs", type=int, default=TRAIN["epochs"])
    t.add_argument("--lr", type=float, default=TRAIN["lr"])
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=WEIGHTS)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--seeds", default="2024,11,12,13")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[proposal] {m}", file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "train":
        data = training_set(a.pairs, a.rotations, log)
        prop = Proposal(max_rot=a.rotations, seed=a.seed)
        train(prop, data, a.epochs, a.lr, a.seed, log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.save(a.out, prop.params())
        log(f"weights -> {a.out}")
    else:
        prop = Proposal(max_rot=30000)
        prop.load(a.weights)
        for s in (int(x) for x in a.seeds.split(",")):
            r = evaluate(prop, seed=s, log=log)
            print(r, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

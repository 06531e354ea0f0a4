"""CPU oracle of the learned rotation proposal (test infrastructure only).

The proposal network has no counterpart in the reference; this restates the
model that include/pdprop.h describes, in FP64 numpy, so that the GPU kernels
(proddl2_amd/csrc/prop_kernels.hpp) can be checked against it:

  maps(pos)            interatomic-vector maps, 2 x 16^3
  conv_stack(...)      3 x (3x3x3 conv), ReLU after the first two, + input
  rotation_functions   RF[r][c*3+s] = sum_{v in shell s} FR[c](v) FL[c](R^T v)
  forward / backward   standardisation, MLP 30 -> 384 -> 384 -> 1, MSE

With ``bf16=True`` every weight read by a convolution or a linear layer is
rounded to BF16 (nearest even), and both operands of the two wide matrix
products (and of the products in their backward pass) are rounded too, as
the MFMA kernels do; with ``bf16=False`` the function is smooth and its
backward is checked against torch autograd in FP64 and finite differences
(tests/test_proposal_oracle.py).

Only tests/ and bench.py may import this module.
"""
from __future__ import annotations

import numpy as np

G, DX, SIGMA, TRUNC, RMAX, RSURF, NSURF = 16, 0.9, 0.5, 2.0, 7.2, 3.3, 12
CIN, C, NSH, HID = 2, 8, 3, 384
NF, NIN = C * NSH, C * NSH + 6
SHELL_R = (2.0, 3.6, 4.8, 7.2)
CW1, CW2 = C * CIN * 27, C * C * 27
CONV_STACK = CW1 + C + CW2 + C + CW2 + C
OFF_W1 = 2 * CONV_STACK
OFF_B1 = OFF_W1 + HID * NIN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAM = OFF_B3 + 1

GRID = (np.arange(G) - 7.5) * DX


def bf16(x) -> np.ndarray:
    """Round to the nearest BF16 (ties to even) through FP32; returns float64."""
    x32 = np.ascontiguousarray(x, np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


# ---------------------------------------------------------------------------
# maps
# ---------------------------------------------------------------------------
def surface_flags(pos: np.ndarray) -> np.ndarray:
    pos = np.asarray(pos, np.float64)
    d = pos[None, :, :] - pos[:, None, :]
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    cnt = (d2 < RSURF * RSURF).sum(1) - 1
    return (cnt < NSURF).astype(np.float64)


def maps(pos: np.ndarray) -> np.ndarray:
    """2 x 16^3 map: sum over ordered pairs i != j with |v| < RMAX of the
    per-axis-truncated Gaussian at v = x_j - x_i, divided by n (channel 1:
    both atoms surface)."""
    pos = np.asarray(pos, np.float64)
    n = pos.shape[0]
    surf = surface_flags(pos)
    d = pos[None, :, :] - pos[:, None, :]  # d[i, j] = x_j - x_i
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    ii, jj = np.nonzero((d2 < RMAX * RMAX) & ~np.eye(n, dtype=bool))
    v = d[ii, jj]
    both = surf[ii] * surf[jj]
    P = np.zeros((CIN, G, G, G))
    lo = np.clip(np.ceil((v - TRUNC) / DX + 7.5).astype(int), 0, G - 1)
    hi = np.clip(np.floor((v + TRUNC) / DX + 7.5).astype(int), 0, G - 1)
    ks = lo[:, :, None] + np.arange(5)[None, None, :]          # pairs, axis, 5
    ok = ks <= hi[:, :, None]
    kc = np.minimum(ks, G - 1)
    dd = (kc - 7.5) * DX - v[:, :, None]
    e = np.where(ok, np.exp(-(dd * dd) / (2.0 * SIGMA * SIGMA)), 0.0)
    for a in range(5):
        for b in range(5):
            for c in range(5):
                w = e[:, 0, a] * e[:, 1, b] * e[:, 2, c]
                m = w != 0.0
                idx = (kc[m, 0, a], kc[m, 1, b], kc[m, 2, c])
                np.add.at(P[0], idx, w[m])
                np.add.at(P[1], idx, (w * both)[m])
    return P / n


# ---------------------------------------------------------------------------
# convolutions
# ---------------------------------------------------------------------------
TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shifted(x: np.ndarray) -> np.ndarray:
    """S[ci, t, i, j, k] = x[ci, i+di, j+dj, k+dk] (zero outside)."""
    cin = x.shape[0]
    xp = np.zeros((cin, G + 2, G + 2, G + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    S = np.empty((cin, 27, G, G, G))
    for t, (di, dj, dk) in enumerate(TAPS):
        S[:, t] = xp[:, 1 + di:1 + di + G, 1 + dj:1 + dj + G, 1 + dk:1 + dk + G]
    return S


def conv(x, w, b, relu, on):
    cout, cin = b.shape[0], x.shape[0]
    W = _r(w, on).reshape(cout, cin, 27)
    y = np.einsum("oct,ctxyz->oxyz", W, _shifted(x)) + np.asarray(b, np.float64)[:, None, None, None]
    return np.maximum(y, 0.0) if relu else y


def conv_bwd(dy, x, w, on):
    """(dx, dw, db) for y = conv(x, w, b) (dy already through the ReLU)."""
    cout, cin = dy.shape[0], x.shape[0]
    W = _r(w, on).reshape(cout, cin, 27)
    S = _shifted(x)
    dw = np.einsum("oxyz,ctxyz->oct", dy, S).reshape(-1)
    db = dy.reshape(cout, -1).sum(1)
    # dx[ci, p] = sum_co sum_t W[co, ci, t] dy[co, p - off(t)]
    dyp = np.zeros((cout, G + 2, G + 2, G + 2))
    dyp[:, 1:-1, 1:-1, 1:-1] = dy
    dx = np.zeros((cin, G, G, G))
    for t, (di, dj, dk) in enumerate(TAPS):
        sh = dyp[:, 1 - di:1 - di + G, 1 - dj:1 - dj + G, 1 - dk:1 - dk + G]
        dx += np.einsum("oc,oxyz->cxyz", W[:, :, t], sh)
    return dx, dw, db


def stack_slices(m: int):
    o = m * CONV_STACK
    sl = {}
    for name, size in (("w1", CW1), ("b1", C), ("w2", CW2), ("b2", C), ("w3", CW2), ("b3", C)):
        sl[name] = slice(o, o + size)
        o += size
    return sl


def conv_stack(P, params, m, on):
    s = stack_slices(m)
    h1 = conv(P, params[s["w1"]], params[s["b1"]], True, on)
    h2 = conv(h1, params[s["w2"]], params[s["b2"]], True, on)
    y3 = conv(h2, params[s["w3"]], params[s["b3"]], False, on)
    F = y3 + P[np.arange(C) % CIN]
    return F, (h1, h2)


# ---------------------------------------------------------------------------
# rotation functions
# ---------------------------------------------------------------------------
def shell_vectors():
    """Listed grid vectors in shell order: (cell indices, vectors, shell ids)."""
    p = np.arange(G ** 3)
    v = np.stack([GRID[p >> 8], GRID[(p >> 4) & 15], GRID[p & 15]], 1)
    r = np.sqrt(v[:, 0] ** 2 + v[:, 1] ** 2 + v[:, 2] ** 2)
    cells, sid = [], []
    for s in range(NSH):
        m = np.nonzero((r >= SHELL_R[s]) & (r < SHELL_R[s + 1]))[0]
        cells.append(m)
        sid.append(np.full(len(m), s))
    cells = np.concatenate(cells)
    return cells, v[cells], np.concatenate(sid)


def _corners(R, vec):
    """Corner cells (n, nv, 8, 3) and weights (n, nv, 8) of q = R^T v."""
    q = np.einsum("rji,vj->rvi", R, vec)
    u = q / DX + 7.5
    i0 = np.floor(u)
    f = u - i0
    i0 = i0.astype(int)
    cells, wts = [], []
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                idx = i0 + np.array([a, b, c])
                w = (f[..., 0] if a else 1 - f[..., 0]) * (f[..., 1] if b else 1 - f[..., 1]) * \
                    (f[..., 2] if c else 1 - f[..., 2])
                ok = np.all((idx >= 0) & (idx < G), -1)
                cells.append(np.where(ok[..., None], idx, 0))
                wts.append(np.where(ok, w, 0.0))
    return np.stack(cells, 2), np.stack(wts, 2)


def rotation_functions(FR, FL, R, chunk=64):
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    cells, vec, sid = shell_vectors()
    FRv = FR.reshape(C, -1)[:, cells]  # C, nv
    out = np.zeros((len(R), C, NSH))
    FLf = FL.reshape(C, -1)
    for s0 in range(0, len(R), chunk):
        cc, ww = _corners(R[s0:s0 + chunk], vec)
        flat = (cc[..., 0] * G + cc[..., 1]) * G + cc[..., 2]
        smp = np.einsum("rvq,crvq->rvc", ww, FLf[:, flat])
        for s in range(NSH):
            m = sid == s
            out[s0:s0 + chunk, :, s] = np.einsum("rvc,cv->rc", smp[:, m], FRv[:, m])
    return out.reshape(len(R), NF)


def rotation_functions_bwd(FR, FL, R, drf, chunk=64):
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    cells, vec, sid = shell_vectors()
    FRv = FR.reshape(C, -1)[:, cells]
    FLf = FL.reshape(C, -1)
    g = np.asarray(drf, np.float64).reshape(len(R), C, NSH)
    dFRv = np.zeros_like(FRv)
    dFL = np.zeros((C, G ** 3))
    for s0 in range(0, len(R), chunk):
        cc, ww = _corners(R[s0:s0 + chunk], vec)
        flat = (cc[..., 0] * G + cc[..., 1]) * G + cc[..., 2]
        smp = np.einsum("rvq,crvq->rvc", ww, FLf[:, flat])
        gv = g[s0:s0 + chunk][:, :, sid]          # r, c, v
        dFRv += np.einsum("rcv,rvc->cv", gv, smp)
        contrib = np.einsum("rcv,cv,rvq->crvq", gv, FRv, ww)
        for c in range(C):
            np.add.at(dFL[c], flat.reshape(-1), contrib[c].reshape(-1))
    dFR = np.zeros((C, G ** 3))
    dFR[:, cells] = dFRv
    return dFR.reshape(C, G, G, G), dFL.reshape(C, G, G, G)


# ---------------------------------------------------------------------------
# full network
# ---------------------------------------------------------------------------
def standardize(x, eps):
    m = x.mean(0)
    sd = np.sqrt(((x - m) ** 2).sum(0) / (x.shape[0] - 1))
    return (x - m) / (sd + eps), m, sd


def standardize_bwd(dz, x, m, sd, eps):
    n = x.shape[0]
    a = 1.0 / (sd + eps)
    b = (dz * (x - m)).sum(0) / ((sd + eps) ** 2 * (n - 1) * sd)
    return a * (dz - dz.mean(0)) - b * (x - m)


def forward(params, PR, PL, R, bf=False):
    """Scores of the rotations R (n x 9) and the cache for backward."""
    params = np.asarray(params, np.float64)
    R = np.asarray(R, np.float64).reshape(-1, 3, 3)
    FR, cr = conv_stack(PR, params, 0, bf)
    FL, cl = conv_stack(PL, params, 1, bf)
    rf = rotation_functions(FR, FL, R)
    z, m, sd = standardize(rf, 1e-6)
    X0 = np.concatenate([z, R[:, :, 0], R[:, :, 1]], 1)
    W1 = params[OFF_W1:OFF_B1].reshape(HID, NIN)
    W2 = params[OFF_W2:OFF_B2].reshape(HID, HID)
    W3 = params[OFF_W3:OFF_B3]
    H1 = np.maximum(_r(X0, bf) @ _r(W1, bf).T + params[OFF_B1:OFF_W2], 0.0)
    H2 = np.maximum(_r(H1, bf) @ _r(W2, bf).T + params[OFF_B2:OFF_W3], 0.0)
    out = H2 @ _r(W3, bf) + params[OFF_B3]
    cache = dict(FR=FR, FL=FL, cr=cr, cl=cl, rf=rf, m=m, sd=sd, X0=X0, H1=H1, H2=H2, R=R, PR=PR, PL=PL)
    return out, cache


def loss_and_grad(params, PR, PL, R, energy, bf=False):
    """MSE against the standardised energies and its gradient (flat, params layout)."""
    params = np.asarray(params, np.float64)
    out, k = forward(params, PR, PL, R, bf)
    e = np.asarray(energy, np.float64)
    y = (e - e.mean()) / np.sqrt(((e - e.mean()) ** 2).sum() / (len(e) - 1))
    n = len(out)
    loss = np.mean((out - y) ** 2)
    dout = 2.0 * (out - y) / n
    g = np.zeros(NPARAM)
    W1 = params[OFF_W1:OFF_B1].reshape(HID, NIN)
    W2 = params[OFF_W2:OFF_B2].reshape(HID, HID)
    W3 = params[OFF_W3:OFF_B3]
    H1, H2, X0 = k["H1"], k["H2"], k["X0"]
    g[OFF_W3:OFF_B3] = dout @ H2
    g[OFF_B3] = dout.sum()
    dZ2 = np.outer(dout, _r(W3, bf)) * (H2 > 0)
    g[OFF_W2:OFF_B2] = (_r(dZ2, bf).T @ _r(H1, bf)).reshape(-1)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ1 = (_r(dZ2, bf) @ _r(W2, bf)) * (H1 > 0)
    g[OFF_W1:OFF_B1] = (_r(dZ1, bf).T @ _r(X0, bf)).reshape(-1)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    dX0 = _r(dZ1, bf) @ _r(W1, bf)
    drf = standardize_bwd(dX0[:, :NF], k["rf"], k["m"], k["sd"], 1e-6)
    dFR, dFL = rotation_functions_bwd(k["FR"], k["FL"], k["R"], drf)
    for mol, dF, P, (h1, h2) in ((0, dFR, PR, k["cr"]), (1, dFL, PL, k["cl"])):
        s = stack_slices(mol)
        dh2, g[s["w3"]], g[s["b3"]] = conv_bwd(dF, h2, params[s["w3"]], bf)
        dy2 = dh2 * (h2 > 0)
        dh1, g[s["w2"]], g[s["b2"]] = conv_bwd(dy2, h1, params[s["w2"]], bf)
        dy1 = dh1 * (h1 > 0)
        _, g[s["w1"]], g[s["b1"]] = conv_bwd(dy1, P, params[s["w1"]], bf)
    return loss, g, out


def select_lowest(scores: np.ndarray, keep: int) -> np.ndarray:
    """Indices (ascending) of the keep lowest scores; ties at the cut go to the lower index."""
    order = np.lexsort((np.arange(len(scores)), np.asarray(scores, np.float32)))
    return np.sort(order[:keep])

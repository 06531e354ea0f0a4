"""Seeded synthetic docking inputs: globular complexes, rotation sets, terms.

Complexes: atoms sit on a randomly oriented, randomly offset FCC lattice with
nearest-neighbour distance 3.0 A (density sqrt(2)/27 = 0.0524 A^-3), cut to a
sphere of radius (3n / (4 pi 0.05))^(1/3) and thinned at random to exactly n
atoms (0.05 A^-3).  A uniform-in-sphere sample with a 3.0 A minimum pair
distance at 0.05 A^-3 is infeasible (hard-sphere packing fraction 0.707), so
the lattice is a stated deviation.  Atom types are uniform over 20 classes and
charges are -1/0/+1 with p = 0.1/0.8/0.1.

Frames follow the reference worker's conventions: the receptor is centred at
the origin, the ligand is moved so that the midpoint of its diameter is at the
origin (include/PRODDL/docking_molstruct.hpp:215-226).

Terms: the reference ships only the LJ pair (potentialName "ljComp",
docking_molforce.hpp:586-602); the shape / electrostatics / desolvation terms
below are new terms with the same structure (an additive radial field
projected from each receptor atom, a per-atom weight deposited for each ligand
atom), so one projection kernel and one correlation pipeline serve all of
them.  Their four ligand weight vectors are linearly independent.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .pdfft import FIELD_COULOMB, FIELD_DESOLV, FIELD_TWOSTEP

N_TYPES = 20
DENSITY = 0.05
NN_DIST = 3.0


@dataclass
class Molecule:
    pos: np.ndarray          # (n, 3) float64
    types: np.ndarray        # (n,) int32 in [0, 20)
    charge: np.ndarray       # (n,) float64 in {-1, 0, 1}


@dataclass
class Complex:
    receptor: Molecule
    ligand: Molecule
    type_desolv: np.ndarray  # (20,) per-type desolvation weights ~ U(0.1, 1)
    type_size: np.ndarray    # (20,) per-type core weights ~ U(0.5, 1.5)
    seed: int = 0
    meta: dict = field(default_factory=dict)


def random_rotation(rng: np.random.Generator, n: int | None = None) -> np.ndarray:
    """Uniform SO(3) samples (Shoemake's uniform quaternions) as 3x3 matrices."""
    m = 1 if n is None else n
    u1, u2, u3 = rng.random(m), rng.random(m), rng.random(m)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R[0] if n is None else R


def rotation_set(n: int, seed: int) -> np.ndarray:
    """n seeded uniform rotations, (n, 9) row-major float64; the first is the identity."""
    rng = np.random.default_rng(seed)
    R = random_rotation(rng, n).reshape(n, 9)
    if n:
        R[0] = np.eye(3).reshape(9)
    return R


def globule(n: int, rng: np.random.Generator) -> np.ndarray:
    radius = (3.0 * n / (4.0 * np.pi * DENSITY)) ** (1.0 / 3.0)
    a = NN_DIST * np.sqrt(2.0)  # cubic FCC lattice constant
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    k = int(np.ceil(radius / a)) + 2
    g = np.arange(-k, k + 1) * a
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = (cells[:, None, :] + basis[None]).reshape(-1, 3)
    pts = pts + rng.random(3) * a
    pts = pts @ random_rotation(rng).T
    inside = pts[np.einsum("ij,ij->i", pts, pts) <= radius * radius]
    if inside.shape[0] < n:
        raise ValueError("lattice sphere holds fewer atoms than requested")
    idx = np.sort(rng.choice(inside.shape[0], size=n, replace=False))
    return np.ascontiguousarray(inside[idx])


def diameter_centre(pos: np.ndarray) -> np.ndarray:
    """Midpoint of the farthest atom pair (exact, by brute force)."""
    best, bi, bj = -1.0, 0, 0
    for s in range(0, pos.shape[0], 512):
        d = ((pos[s:s + 512, None, :] - pos[None, :, :]) ** 2).sum(-1)
        k = int(np.argmax(d))
        i, j = divmod(k, pos.shape[0])
        if d[i, j] > best:
            best, bi, bj = float(d[i, j]), s + i, j
    return 0.5 * (pos[bi] + pos[bj])


def molecule(n: int, rng: np.random.Generator) -> Molecule:
    pos = globule(n, rng)
    types = rng.integers(0, N_TYPES, size=n).astype(np.int32)
    charge = rng.choice(np.array([-1.0, 0.0, 1.0]), size=n, p=[0.1, 0.8, 0.1])
    return Molecule(pos, types, charge)


def make_complex(n_rec: int, n_lig: int, seed: int) -> Complex:
    rng = np.random.default_rng(seed)
    rec = molecule(n_rec, rng)
    lig = molecule(n_lig, rng)
    rec.pos = rec.pos - rec.pos.mean(0)
    lig.pos = lig.pos - diameter_centre(lig.pos)
    type_desolv = rng.uniform(0.1, 1.0, N_TYPES)
    type_size = rng.uniform(0.5, 1.5, N_TYPES)
    return Complex(rec, lig, type_desolv, type_size, seed)


# ---------------------------------------------------------------------------
# Terms.  Each term: receptor (kind, cutoff, min_dist, params[n_rec, np]) and a
# ligand weight vector.
# ---------------------------------------------------------------------------
CORE_RADIUS = 3.0
CORE_PENALTY = 3.0
SURFACE_RADIUS = 5.0
COULOMB_CUTOFF = 9.0
COULOMB_MIN_DIST = 2.0
DESOLV_CUTOFF = 6.0
COULOMB_K = 332.0 / 4.0  # 332 q q' / (eps r), eps = 4 r


def terms_shape_elec(cx: Complex):
    """Two terms: combined shape (core clash + surface contact) and Coulomb."""
    r, l = cx.receptor, cx.ligand
    n = r.pos.shape[0]
    shape = np.tile([CORE_PENALTY, -1.0, CORE_RADIUS ** 2], (n, 1))
    elec = (COULOMB_K * r.charge).reshape(n, 1)
    rec_terms = [
        (FIELD_TWOSTEP, SURFACE_RADIUS, -1.0, shape),
        (FIELD_COULOMB, COULOMB_CUTOFF, COULOMB_MIN_DIST, elec),
    ]
    w = np.stack([np.ones(l.pos.shape[0]), l.charge])
    return rec_terms, w


def terms_four(cx: Complex):
    """Four terms: shape core, shape surface, electrostatics, desolvation proxy."""
    r, l = cx.receptor, cx.ligand
    n = r.pos.shape[0]
    core = np.tile([CORE_PENALTY, 0.0, CORE_RADIUS ** 2], (n, 1))
    surf = np.tile([0.0, -1.0, CORE_RADIUS ** 2], (n, 1))
    elec = (COULOMB_K * r.charge).reshape(n, 1)
    desolv = cx.type_desolv[r.types].reshape(n, 1)
    rec_terms = [
        (FIELD_TWOSTEP, CORE_RADIUS, -1.0, core),
        (FIELD_TWOSTEP, SURFACE_RADIUS, -1.0, surf),
        (FIELD_COULOMB, COULOMB_CUTOFF, COULOMB_MIN_DIST, elec),
        (FIELD_DESOLV, DESOLV_CUTOFF, -1.0, desolv),
    ]
    w = np.stack([
        cx.type_size[l.types],
        np.ones(l.pos.shape[0]),
        l.charge,
        cx.type_desolv[l.types],
    ])
    return rec_terms, w


def terms_for(cx: Complex, n_terms: int):
    if n_terms == 2:
        return terms_shape_elec(cx)
    if n_terms == 4:
        return terms_four(cx)
    raise ValueError("synthetic complexes define 2 or 4 terms")


# Lennard-Jones parameters per atom type, in the units the reference's molforce
# writer produces (python/lib/PRODDL/force_field_fft.py:91-99): sigma =
# 2 * 2^(-1/6) * (r_min / 2) with r_min / 2 of 1.7-2.2 A, eps = 4.184 * |eps_charmm|
# with |eps_charmm| of 0.02-0.2 kcal/mol; alpha is the scan's soft-core
# parameter (default 0.4, config/noarch/proddl.json.in:11).
LJ_SIGMA_RANGE = (3.03, 3.92)
LJ_EPS_RANGE = (0.084, 0.84)
LJ_CUTOFF = 9.0  # cutOffFft default (config/noarch/proddl.json.in:14)


@dataclass
class LJParams:
    mass: np.ndarray   # (n,) per atom of receptor then ligand
    eps: np.ndarray
    sigma: np.ndarray
    alpha: np.ndarray
    n_rec: int


def lj_params(cx: Complex, alpha=0.4, seed: int | None = None) -> LJParams:
    """Per-atom mass, eps, sigma and alpha of receptor then ligand atoms, drawn per
    atom type.  alpha: a number, or (lo, hi) for per-atom U(lo, hi) values."""
    rng = np.random.default_rng(cx.seed + 7919 if seed is None else seed)
    sig_t = rng.uniform(*LJ_SIGMA_RANGE, N_TYPES)
    eps_t = rng.uniform(*LJ_EPS_RANGE, N_TYPES)
    mass_t = rng.uniform(12.0, 32.0, N_TYPES)
    types = np.concatenate([cx.receptor.types, cx.ligand.types])
    n = types.shape[0]
    if np.ndim(alpha) == 0:
        al = np.full(n, float(alpha))
    else:
        al = rng.uniform(float(alpha[0]), float(alpha[1]), n)
    return LJParams(mass_t[types], eps_t[types], sig_t[types], al, cx.receptor.pos.shape[0])


def terms_ljcomp(cx: Complex, alpha=0.4, cutoff: float = LJ_CUTOFF, seed: int | None = None):
    """The reference's own two-term potential, potentialName "ljComp"
    (MolForceLJComp, docking_molforce.hpp:586-602): soft-core LJ repulsion and
    attraction fields of the receptor (PD_FIELD_LJ_REP / LJ_ATT) with the ligand
    weights sqrt(eps) sigma^6 and sqrt(eps) sigma^3, all from pd_ljcomp_params."""
    from .pdfft import FIELD_LJ_ATT, FIELD_LJ_REP, ljcomp_params
    p = lj_params(cx, alpha, seed)
    nr = p.n_rec
    rep, att, wr, wa = ljcomp_params(p.eps[:nr], p.sigma[:nr], p.alpha[:nr], p.eps[nr:], p.sigma[nr:])
    rec_terms = [(FIELD_LJ_REP, cutoff, -1.0, rep), (FIELD_LJ_ATT, cutoff, -1.0, att)]
    return rec_terms, np.stack([wr, wa])


# Workloads, by grid and sizes.
WORKLOADS = {
    "small-64": dict(n_rec=1800, n_lig=450, n=(64, 64, 64), h=1.2, n_terms=2, n_rot=500, K=10),
    "exhaustive-128": dict(n_rec=4000, n_lig=1500, n=(128, 128, 128), h=1.2, n_terms=4,
                           n_rot=30000, K=10),
    "stress-256": dict(n_rec=12000, n_lig=3000, n=(256, 256, 256), h=1.0, n_terms=4,
                       n_rot=10000, K=10),
}


def terms_ljavg(cx: Complex, alpha=0.4, cutoff: float = LJ_CUTOFF, mix: int = 1, seed: int | None = None):
    """The reference's second potential, potentialName "ljAvg" (MolForceLJ,
    docking_molforce.hpp:377-450): one term, the full soft-core LJ with a
    shifted-force cutoff of each receptor atom mixed with the ligand's average
    atom (PD_FIELD_LJ_FULL), ligand weights 1; parameters from pd_ljavg_params."""
    from .pdfft import FIELD_LJ_FULL, ljavg_params
    p = lj_params(cx, alpha, seed)
    nr = p.n_rec
    par, w = ljavg_params(p.eps[:nr], p.sigma[:nr], p.alpha[:nr], p.eps[nr:], p.sigma[nr:], mix, cutoff)
    return [(FIELD_LJ_FULL, cutoff, -1.0, par)], w.reshape(1, -1)

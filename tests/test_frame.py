"""The docking frame's geometry (host code of the library, no device calls):
pd_min_volume_box against the reference's minimum-volume box
(MolStruct::moveIntoMinBox, docking_molstruct.hpp:199-245, which calls the
vendored gdiam through Geom::Bounding::Box, Geom/bounding.hpp:135-183) and
pd_diameter_pair against Geom::Bounding::Diameter (bounding.hpp:33-97).

Pins:
  * closed forms: a rotated cuboid given by its corners and interior points is
    its own minimum box;
  * brute force: no box among 20,000 random orientations of a small point set is
    smaller;
  * the reference's own gdiam on its own test complex (2ptc, golden fixture
    written by tests/golden/make_frame_golden.py): the same box up to the order
    and sign of its axes, the same diameter pairs;
  * gdiam built from the reference sources (oracle/_ref), where present, on
    further point sets.
"""
import os

import numpy as np
import pytest

from proddl2_amd import pdfft, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frame_2ptc.npz")


def _check_box(pts, axes, lo, hi):
    assert np.abs(axes @ axes.T - np.eye(3)).max() < 1e-12
    assert abs(np.linalg.det(axes) - 1.0) < 1e-12
    proj = pts @ axes.T
    assert np.allclose(proj.min(0), lo, rtol=0, atol=1e-9) and np.allclose(proj.max(0), hi, rtol=0, atol=1e-9)
    ext = hi - lo
    assert np.all(np.diff(ext) <= 1e-9), ext  # rows by decreasing extent


def _axes_match(a, b, tol):
    """Rows of a equal rows of b up to order and sign."""
    m = np.abs(a @ b.T)
    return np.all(np.abs(np.sort(m, axis=1)[:, -1] - 1.0) < tol) and sorted(np.argmax(m, axis=1)) == [0, 1, 2]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rotated_cuboid_is_its_own_box(seed):
    rng = np.random.default_rng(seed)
    ext = np.array([31.0, 17.5, 9.25]) * rng.uniform(0.8, 1.2, 3)
    ext = np.sort(ext)[::-1]
    R = synth.random_rotation(rng)
    corners = np.array([[x, y, z] for x in (-.5, .5) for y in (-.5, .5) for z in (-.5, .5)]) * ext
    pts = np.concatenate([corners, rng.uniform(-.5, .5, (400, 3)) * ext]) @ R.T + rng.normal(0, 20, 3)
    axes, lo, hi = pdfft.min_volume_box(pts)
    _check_box(pts, axes, lo, hi)
    assert np.allclose(hi - lo, ext, rtol=0, atol=1e-9)
    # the cuboid's axes (columns of R), each up to sign
    assert np.allclose(np.abs(axes @ R), np.eye(3), atol=1e-9)


def test_axis_parallel_cuboid_gives_identity_axes():
    rng = np.random.default_rng(7)
    ext = np.array([40.0, 30.0, 20.0])
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)]) * ext
    pts = np.concatenate([corners, rng.uniform(0, 1, (200, 3)) * ext]) + np.array([-15.0, -10.0, -5.0])
    axes, lo, hi = pdfft.min_volume_box(pts)
    assert np.abs(axes - np.eye(3)).max() < 1e-12
    assert np.allclose(lo, [-15, -10, -5], atol=1e-12) and np.allclose(hi, [25, 20, 15], atol=1e-12)


@pytest.mark.parametrize("seed", [3, 4])
def test_no_sampled_orientation_is_smaller(seed):
    rng = np.random.default_rng(seed)
    pts = rng.normal(0, 1, (40, 3)) * np.array([9.0, 5.0, 3.0])
    axes, lo, hi = pdfft.min_volume_box(pts)
    _check_box(pts, axes, lo, hi)
    vol = np.prod(hi - lo)
    R = synth.random_rotation(rng, 20000)
    proj = np.einsum("rij,nj->rni", R, pts)
    sampled = np.prod(proj.max(1) - proj.min(1), axis=1)
    assert vol <= sampled.min() * (1 + 1e-9), (vol, sampled.min())
    assert vol <= np.prod(np.ptp(pts, 0))


def test_diameter_pair_is_exact():
    rng = np.random.default_rng(5)
    for n in (2, 3, 57, 300):
        pts = rng.normal(0, 10, (n, 3))
        i, j = pdfft.diameter_pair(pts)
        d2 = ((pts[:, None] - pts[None]) ** 2).sum(-1)
        assert i < j and d2[i, j] == d2.max()


def test_invalid_points_are_refused():
    with pytest.raises(pdfft.PdError):
        pdfft.min_volume_box(np.array([[0.0, np.nan, 1.0]]))
    with pytest.raises(pdfft.PdError):
        pdfft.diameter_pair(np.zeros((0, 3)))


def test_matches_reference_gdiam_on_2ptc():
    """The reference's test complex: the receptor's box equals gdiam's (axes up
    to order and sign to 1e-3 rad, extents to 0.01 A, volume to 1e-4); the
    diameter pairs are the same atoms."""
    g = np.load(GOLDEN)
    rec, lig = g["receptor"], g["ligand"]
    axes, lo, hi = pdfft.min_volume_box(rec)
    _check_box(rec, axes, lo, hi)
    dirs, left, right = g["box_dirs"], g["box_left"], g["box_right"]
    gext = np.abs(dirs @ (right - left))
    assert _axes_match(axes, dirs, 1e-6)
    assert np.allclose(np.sort(hi - lo), np.sort(gext), atol=0.01)
    assert abs(np.prod(hi - lo) / np.prod(gext) - 1) < 1e-4
    for pts, pair in ((rec, g["rec_diam"]), (lig, g["lig_diam"])):
        i, j = pdfft.diameter_pair(pts)
        assert {tuple(pts[i]), tuple(pts[j])} == {tuple(pair[0]), tuple(pair[1])}


def test_grid_edges_follow_reference_rule_on_2ptc(oracle):
    """The padded box's grid edges (FFTCorrelator::init through findBestSize,
    docking_fft.hpp:541-553) from this box and from gdiam's are the same
    multiset: the frames differ at most by the order of the axes."""
    g = np.load(GOLDEN)
    rec, lig = g["receptor"], g["ligand"]
    axes, lo, hi = pdfft.min_volume_box(rec)
    gext = np.abs(g["box_dirs"] @ (g["box_right"] - g["box_left"]))
    lig_size = np.linalg.norm(g["lig_diam"][0] - g["lig_diam"][1])
    for h in (0.8, 1.0, 1.2, 1.5):
        pad = 2 * h + lig_size / 2 + 9.0
        mine = sorted(oracle.grid_size_from_box(-e / 2 - pad, e / 2 + pad, h) for e in hi - lo)
        ref = sorted(oracle.grid_size_from_box(-e / 2 - pad, e / 2 + pad, h) for e in gext)
        assert mine == ref, (h, mine, ref)


def test_matches_reference_gdiam_built_from_source():
    """gdiam compiled from the reference's own sources (oracle/_ref): the same
    box on point sets of 400-6,000 atoms.  The synthetic lattice is jittered by
    0.05 A: gdiam asserts on exactly collinear hull points."""
    from oracle import gdiam_ref
    if gdiam_ref.load() is None:
        pytest.skip("oracle/_ref/libgdiam_ref.so not built (reference sources absent)")
    rng = np.random.default_rng(11)
    for n, seed in ((400, 1), (1500, 2), (6000, 3)):
        p = synth.make_complex(n, 100, seed).receptor.pos
        p = p + rng.normal(0, 0.05, p.shape)
        axes, lo, hi = pdfft.min_volume_box(p)
        dirs, left, right = gdiam_ref.mvbb(p)
        gext = np.abs(dirs @ (right - left))
        ratio = np.prod(hi - lo) / np.prod(gext)
        print(f"  {n} atoms: volume ratio to gdiam {ratio:.6f}", flush=True)
        assert ratio < 1 + 1e-4
        assert _axes_match(axes, dirs, 1e-3) or ratio < 1 - 1e-4

"""Pins of the FP64 CPU oracle (oracle/pdoracle.cpp).

Each oracle function is checked against (a) a known answer restated from the
reference's own tests, (b) an independent library routine (numpy's pocketfft),
or (c) a closed form / brute force that a plausible mistake would fail.
The reference itself cannot be built here (Blitz++, FFTW3, Boost and HDF5
headers are absent), so these are the pins.
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# -- Math/test_pfactors.cpp:73-132: sieve vs findFirstPrimeFactoring / findBestSize
def test_pfactors_sieve(oracle):
    primes = [2, 3, 5, 7, 11, 13]
    nmax = 1024

    def factor(n):
        e = []
        for p in primes:
            k = 0
            while n % p == 0:
                n //= p
                k += 1
            e.append(k)
        return n == 1, e

    smooth = [n for n in range(nmax) if n >= 2 and factor(n)[0]]
    best = [n for n in range(2, 2 * nmax) if factor(n)[0] and factor(n)[1][4] + factor(n)[1][5] <= 1]
    for i in range(nmax):
        nxt = next((n for n in smooth if n >= max(i, 2)), None)
        ok, found, ex = oracle.find_first_prime_factoring(i, nmax, primes)
        if nxt is None:
            assert not ok
            continue
        assert ok and found == nxt
        assert int(np.prod([p ** int(k) for p, k in zip(primes, ex)])) == found
        nb = next(n for n in best if n >= max(i, 2))
        assert oracle.find_best_size(i) == nb


def test_find_best_size_library_matches_oracle(oracle):
    from proddl2_amd import pdfft
    for i in list(range(0, 300)) + [511, 513, 1000, 1025, 2047]:
        assert pdfft.find_best_size(i) == oracle.find_best_size(i)


# -- Common/test_queue.cpp:21-99 restated on the selection path
def test_bounded_queue_known_answer(oracle):
    vals = np.array([1, 5, 11, 12, 4, 3, -5, 2], float).reshape(1, 1, 8)
    off, s, tran, v = oracle.select(vals, 1.0, 4, 10.0, 4)
    assert list(v) == [-5, 1, 2, 3]
    vals = np.array([5, 4, 12, 1, 2], float).reshape(1, 1, 5)  # back_inserter part
    off, s, tran, v = oracle.select(vals, 1.0, 4, 10.0, 10)
    assert list(v) == [1, 2, 4, 5]


def test_queue_tie_rule_and_row_major_order(oracle):
    # Equal values: a later entry never replaces an equal top (queue.hpp:233-239),
    # so the first max_inp cells in row-major order (nd_index_iter.hpp:55-84) survive.
    g = np.zeros((3, 4, 5))
    off, s, tran, v = oracle.select(g, 1.0, 3, 1.0, 3)
    assert sorted(off) == [0, 1, 2]
    # -0.0 and +0.0 are equal under operator<: the first cells still win
    g = np.zeros((3, 4, 5))
    g.reshape(-1)[[1, 4, 7, 30, 59]] = -0.0
    off, s, tran, v = oracle.select(g, 1.0, 3, 1.0, 3)
    assert sorted(off) == [0, 1, 2]
    # maxValCorr is a strict bound (queue.hpp:225-226)
    g = np.arange(60, dtype=float).reshape(3, 4, 5)
    off, s, tran, v = oracle.select(g, 1.0, 100, 10.0, 100)
    assert list(v) == list(range(10))


def test_select_matches_brute_force_sort(oracle):
    rng = np.random.default_rng(3)
    for shape in [(5, 7, 9), (6, 8, 10), (4, 4, 4)]:
        g = rng.standard_normal(shape)
        K = 12
        off, s, tran, v = oracle.select(g, 1.5, 50, 1e16, K)
        order = np.argsort(g.reshape(-1), kind="stable")[:K]
        assert list(off) == list(order)
        # offset -> displacement: s = -x wrapped into [-floor(N/2), N-floor(N/2)-1]
        x = np.stack(np.unravel_index(order, shape), -1)
        n = np.array(shape)
        exp = -x
        exp = np.where(exp < -(n // 2), exp + n, exp)
        assert (s == exp).all()
        assert np.allclose(tran, 1.5 * exp)


# -- Grid/test_grid.cpp:60-76: centred geometry, step 20, size (50, 51, 51)
def test_centred_geometry_known_answer(oracle):
    for n, half in [(50, 500.0), (51, 510.0)]:
        zero_h, zsc = oracle.geometry_centered(20.0, n)
        assert zero_h + 20.0 == -half          # spatial domain lower bound
        assert zsc + 20.0 * (n - 1) + 10.0 == half  # right edge of the last cell
        assert oracle.to_logical(20.0, zero_h, -half + 1e-9) == 0
        assert oracle.to_logical(20.0, zero_h, half - 1e-9) == n - 1
        assert oracle.to_logical(20.0, zero_h, -half - 1e-9) == -1


def test_cell_of_cell_centre_is_itself(oracle):
    for h, n in [(1.2, 128), (1.5, 75), (1.0, 64)]:
        zero_h, zsc = oracle.geometry_centered(h, n)
        for i in range(n):
            assert oracle.to_logical(h, zero_h, zsc + h * i) == i


def test_grid_size_from_box(oracle):
    # FFTCorrelator::init: N = findBestSize(toLogical(upper corner)) in a geometry
    # whose zero is the lower corner; a box of exactly N*h + eps gives N.
    for n, h in [(64, 1.2), (128, 1.2), (96, 1.0)]:
        half = n * h / 2 + 1e-6
        assert oracle.grid_size_from_box(-half, half, h) == n


# -- Geom/test_transformation.cpp:20-50: composition and inverse known answers
def test_transformation_known_answer(oracle):
    m = np.array([[0, 1, 1], [0.2, 0.3, 0.4], [1, 0.3, 0.5]])
    m2 = np.array([[0.4, 1.7, 1.8], [0.7, 0.9, 0.3], [1.3, 0.1, 0.9]])
    cm = np.array([5.0, 4, 3])
    v = np.array([1.0, 2, 3])
    # rt = Rot(m2) * Tran(cm + v) * (Rot(m) * Tran(-cm))   (transformation.hpp:191-203, 318-327)
    inner_t = oracle.rotate_points(m, (-cm)[None])[0]
    outer_t = oracle.rotate_points(m2, (cm + v)[None])[0]
    T = oracle.matmul3(m2, m)
    t = oracle.rotate_points(m2, inner_t[None])[0] + outer_t
    Ti = T.T                                   # inverse(): Rotation(transpose) * Translation(-t)
    ti = oracle.rotate_points(Ti, (-t)[None])[0]
    exp_T = np.array([[2.14, 0.48, 0.92], [1.45, 1.06, 1.6], [1.98, 1.21, 1.79]])
    exp_t = np.array([-0.2324, 1.5222, 1.3322])
    assert np.allclose(Ti, exp_T, atol=1e-9)
    assert np.allclose(ti, exp_t, atol=1e-9)


# -- Geom/euler.hpp:38-102
def test_euler_closed_forms(oracle):
    assert np.allclose(oracle.euler_to_matrix([0, 0, 0]), np.eye(3))
    rng = np.random.default_rng(1)
    for _ in range(50):
        a = rng.uniform(-np.pi, np.pi, 3)
        a[2] = rng.uniform(-np.pi / 2 + 0.01, np.pi / 2 - 0.01)
        M = oracle.euler_to_matrix(a)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(M) - 1) < 1e-13
        # single-axis rotations: pure gamma rotates about y (m = [[cg,0,-sg],[0,1,0],[sg,0,cg]])
    g = 0.3
    M = oracle.euler_to_matrix([0, 0, g])
    assert np.allclose(M, [[np.cos(g), 0, -np.sin(g)], [0, 1, 0], [np.sin(g), 0, np.cos(g)]])
    # matrix -> angles -> matrix round trip away from gimbal lock
    for _ in range(50):
        a = rng.uniform(-1.2, 1.2, 3)
        M = oracle.euler_to_matrix(a)
        b, unique = oracle.matrix_to_euler(M)
        assert unique
        assert np.allclose(oracle.euler_to_matrix(b), M, atol=1e-12)


def test_rotational_grid_file(oracle):
    path = os.path.join(GOLDEN, "ang30.dat")
    R = oracle.load_rotational_grid(path)
    assert R.shape == (622, 9)  # header count of data/rotgrid/ang30.dat
    assert np.allclose(R[0].reshape(3, 3), np.eye(3))
    a = np.deg2rad([30.0, 0, 0])
    assert np.allclose(R[1], oracle.euler_to_matrix(a).reshape(9))
    for r in R[::37]:
        M = r.reshape(3, 3)
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-13)


# -- FFT: own transform vs numpy's pocketfft
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 11, 13, 16, 60, 64, 77, 91, 96, 120, 128, 143, 97])
def test_fft1_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft1(x, -1), np.fft.fft(x), atol=1e-11 * n)
    assert np.allclose(oracle.fft1(x, +1), np.fft.ifft(x) * n, atol=1e-11 * n)


@pytest.mark.parametrize("shape", [(5, 7, 9), (6, 8, 10), (4, 6, 12), (16, 16, 16), (8, 6, 7)])
def test_rfft3_irfft3_match_numpy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    g = rng.standard_normal(shape)
    S = oracle.rfft3(g)
    assert np.allclose(S, np.fft.rfftn(g), atol=1e-10)
    assert np.allclose(oracle.irfft3(S, shape), g * g.size, atol=1e-9)   # testIdentity, docking_fft.hpp:690-697


def test_irfft3_ignores_self_conjugate_imaginary_parts(oracle):
    # FFTW's c2r drops Im of the bins it treats as real along the last axis.
    shape = (4, 6, 8)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(shape)
    S = np.fft.rfftn(g)
    S2 = S.copy()
    S2[:, :, 0] += 1j * 0  # already hermitian-consistent
    S2[0, 0, 0] += 3j
    S2[0, 0, -1] += 2j
    assert np.allclose(oracle.irfft3(S2, shape), oracle.irfft3(S, shape), atol=1e-9)

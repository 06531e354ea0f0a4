"""pd_scan_rescore64_noalloc (the allocation-free rescoring entry point) against
pd_scan_rescore64 through the C ABI: the same energies, bit for bit, on two calls
in a row (the second smaller) and from a second handle on the same device; more
poses than the staging area holds is an invalid argument."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no GPU", allow_module_level=True)

from proddl2_amd import pdfft, synth  # noqa: E402

MAX_POSES = 8192


def _energies(fn, sc, rots, s):
    r = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
    sv = np.ascontiguousarray(s, np.int32).reshape(-1, 3)
    out = np.full(r.shape[0], np.nan)
    st = fn(sc.handle, r.shape[0], r.ctypes.data_as(C.POINTER(C.c_double)), sv.ctypes.data_as(C.POINTER(C.c_int)),
            out.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out


def _scan(n, h, T, seed):
    cx = synth.make_complex(700, 220, seed)
    rt, w = synth.terms_for(cx, T)
    sc = pdfft.Scan(n, h, T, max_n_trans=4, max_val_corr=1e16)
    sc.project_receptor(cx.receptor.pos, rt)
    sc.set_ligand(cx.ligand.pos, w)
    return sc


def test_noalloc_rescoring_matches_allocating_entry_point():
    lib = pdfft.load()
    rng = np.random.default_rng(2)
    print("case no-alloc rescoring: 48^3 T=2, calls of 3000 then 40 poses; second handle 40^3 T=4", flush=True)
    with _scan((48, 48, 48), 1.5, 2, 51) as sc, _scan((40, 40, 40), 1.5, 4, 52) as sc2:
        for handle, n in ((sc, 3000), (sc, 40), (sc2, 700), (sc, MAX_POSES)):
            rots = synth.rotation_set(n, seed=int(rng.integers(1 << 30)))
            s = rng.integers(-12, 12, size=(n, 3)).astype(np.int32)
            st_a, e_a = _energies(lib.pd_scan_rescore64, handle, rots, s)
            st_b, e_b = _energies(lib.pd_scan_rescore64_noalloc, handle, rots, s)
            assert st_a == 0 and st_b == 0, lib.pd_last_error()
            assert np.array_equal(e_a, e_b), (n, np.abs(e_a - e_b).max())
            assert np.isfinite(e_b).all()
            print(f"  {n} poses: identical", flush=True)
        big = MAX_POSES + 1
        st, e = _energies(lib.pd_scan_rescore64_noalloc, sc, synth.rotation_set(big, seed=3),
                          np.zeros((big, 3), np.int32))
        assert st == 1 and b"at most 8192" in lib.pd_last_error()
        assert np.isnan(e).all()  # nothing written
        st, _ = _energies(lib.pd_scan_rescore64_noalloc, sc, np.zeros((0, 9)), np.zeros((0, 3), np.int32))
        assert st == 0

"""The allocation-free solve entry point against the allocating one.

milc_hip_dev_invert_noalloc takes its work vectors from the library's static
scratch array; milc_hip_dev_invert allocates them.  Both run the same solver
kernels in the same order, so on the same links and source their solutions,
iteration counts and residuals must agree (bar: 1e-14 relative on the
solution, equal iteration counts).  Two calls in a row through the C ABI, the
second on a smaller lattice (the scratch then holds the first call's data),
each in FP64 and in 16-bit mixed precision.  Also: the size limit and the
preconditions are reported as errors, not allocations.
"""
import ctypes

import numpy as np
import pytest

import milc_qcd_amd as mq
from milc_qcd_amd import ks

pytestmark = pytest.mark.gpu


def _noalloc(G, slot_b, slot_x, mass, parity, resid, mixed, max_iter=2000):
    st = (ctypes.c_double * 6)()
    rc = G.lib.milc_hip_dev_invert_noalloc(slot_b, slot_x, ctypes.c_double(mass), parity,
                                           ctypes.c_double(resid), mixed, max_iter, st)
    return rc, {"resid": st[0], "converged": int(st[3]), "iters": int(st[4]), "final_rsq": st[5]}


def _compare_on(dims, seed):
    print(f"noalloc vs allocating solve on {dims}", flush=True)
    with mq.Lattice(dims) as G:
        lib = G.lib
        lib.milc_hip_dev_invert_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_double)]
        G.dev_gauge_hisq(seed, 0.4)
        G.dev_source(seed + 1, mq.EVEN, 0)
        # slot 2 does not exist yet: the allocation-free call must refuse it
        rc, _ = _noalloc(G, 0, 2, 0.05, mq.EVEN, 1e-10, 0)
        assert rc == -3
        for mixed in (0, 2):
            it_a, st_a = G.dev_invert(0, 1, 0.05, mq.EVEN, 1e-10, mixed)
            # give slot 2 stale contents: the solve must zero it itself
            G.dev_write(2, mq.EVEN, np.full((G.V, 3, 2), 7.0))
            rc, st_b = _noalloc(G, 0, 2, 0.05, mq.EVEN, 1e-10, mixed)
            assert rc == it_a == st_b["iters"], (rc, it_a)
            assert st_b["converged"] == 1 and st_a["converged"] == 1
            a = G.dev_read(1, mq.EVEN)
            b = G.dev_read(2, mq.EVEN)
            rel = np.linalg.norm(a - b) / np.linalg.norm(a)
            print(f"  mixed={mixed}: {it_a} iterations, |r|/|b| {st_b['resid']:.3e}, rel diff {rel:.2e}", flush=True)
            assert rel <= 1e-14
            assert st_b["resid"] < 1e-10
        # invalid parity is an error
        rc, _ = _noalloc(G, 0, 2, 0.05, 3, 1e-10, 0)
        assert rc == -2


def test_noalloc_matches_allocating_solve_two_calls():
    _compare_on((8, 8, 8, 16), 1201)
    # second call: a smaller lattice, the scratch still holds the first call's fields
    _compare_on((4, 4, 4, 8), 1301)


def test_noalloc_rejects_lattice_beyond_scratch():
    # 48^3 x 64: 3.5M sites, 200 B per site of one parity in mixed precision
    # exceeds the 256 MiB scratch; FP64 needs 192 B per site and also exceeds it
    dims = (48, 48, 48, 64)
    with mq.Lattice(dims) as G:
        lib = G.lib
        lib.milc_hip_dev_invert_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_double)]
        G.dev_gauge_hisq(7, 0.4)
        G.dev_source(8, mq.EVEN, 0)
        G.dev_source(9, mq.EVEN, 1)
        rc, _ = _noalloc(G, 0, 1, 0.05, mq.EVEN, 1e-6, 0)
        assert rc == -6
    assert ks.load_library() is not None

"""ctypes front end of the compiled MILC reference (oracle/_ref/libmilc_ref.so).

TEST INFRASTRUCTURE ONLY.  The library is MILC 7.8.1's own CPU code compiled
from the reference sources by oracle/Makefile.ref (see ref_driver.c).  MILC
keeps its lattice in process globals, so one process can use one lattice
size: the first init() fixes it.  Fields use MILC's layout (same as
ks_oracle.py): vectors (V, 3, 2), links (V, 4, 3, 3, 2), even sites first.
"""
import ctypes
import os

import numpy as np

EVEN, ODD, EVENANDODD = 2, 1, 3
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libmilc_ref.so")
_lib = None
_dims = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle -f Makefile.ref)")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.ref_init.argtypes = [I, I, I, I]
        L.ref_sites.restype = ctypes.c_size_t
        L.ref_set_thin_links.argtypes = [P]
        L.ref_hisq_links.argtypes = [P, P]
        L.ref_set_fn_links.argtypes = [P, P]
        L.ref_dslash.argtypes = [P, P, I]
        L.ref_congrad.argtypes = [P, P, I, D, I, I, D, ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I)]
        L.ref_congrad_rel.argtypes = [P, P, I, D, I, I, D, D, ctypes.POINTER(D), ctypes.POINTER(D),
                                      ctypes.POINTER(I)]
        L.ref_multicg.argtypes = [P, ctypes.POINTER(P), P, I, I, I, D, ctypes.POINTER(D)]
        L.ref_force_oprod.argtypes = [ctypes.POINTER(P), P, I, P, P]
        L.ref_hisq_force.argtypes = [ctypes.POINTER(P), P, I, ctypes.c_double, P]
        L.ref_hisq_links_naik.argtypes = [D, P, P, P, P]
        L.ref_hisq_force_naik.argtypes = [ctypes.POINTER(P), P, I, I, ctypes.c_double, P]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def init(dims):
    """Set up MILC's lattice once per process; later calls must use the same dims."""
    global _dims
    dims = tuple(int(d) for d in dims)
    if _dims is None:
        if lib().ref_init(*dims) != 0:
            raise RuntimeError("ref_init failed")
        _dims = dims
    elif _dims != dims:
        raise RuntimeError(f"MILC reference already initialised for {_dims}")
    return int(lib().ref_sites())


def set_thin_links(U_unphased):
    lib().ref_set_thin_links(_p(np.ascontiguousarray(U_unphased, dtype=np.float64)))


def hisq_links(V):
    fat = np.zeros((V, 4, 3, 3, 2))
    lng = np.zeros((V, 4, 3, 3, 2))
    lib().ref_hisq_links(_p(fat), _p(lng))
    return fat, lng


def hisq_links_naik(V, eps):
    """MILC's links with two Naik sets, eps_naik = {0, eps}: (fat0, lng0, fat1, lng1)"""
    out = [np.zeros((V, 4, 3, 3, 2)) for _ in range(4)]
    lib().ref_hisq_links_naik(eps, *[_p(a) for a in out])
    return tuple(out)


def set_fn_links(fat, lng):
    lib().ref_set_fn_links(_p(np.ascontiguousarray(fat)), _p(np.ascontiguousarray(lng)))


def dslash(src, parity):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros_like(src)
    lib().ref_dslash(_p(src), _p(dst), parity)
    return dst


def congrad(src, mass, resid, parity=EVEN, niter=1000, nrestart=5, dest=None):
    src = np.ascontiguousarray(src, dtype=np.float64)
    if dest is None:
        dest = np.zeros_like(src)
    frsq, rs, conv = ctypes.c_double(0), ctypes.c_int(0), ctypes.c_int(0)
    it = lib().ref_congrad(_p(src), _p(dest), parity, mass, niter, nrestart, resid, ctypes.byref(frsq),
                           ctypes.byref(rs), ctypes.byref(conv))
    return dest, it, {"final_rsq": frsq.value, "restarts": rs.value, "converged": conv.value}


def congrad_rel(src, mass, resid, relresid, parity=EVEN, niter=1000, nrestart=5):
    """ks_congrad_parity_cpu with the Fermilab relative residual test (qic.relresid)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    dest = np.zeros_like(src)
    frsq, frel, conv = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
    it = lib().ref_congrad_rel(_p(src), _p(dest), parity, mass, niter, nrestart, resid, relresid, ctypes.byref(frsq),
                               ctypes.byref(frel), ctypes.byref(conv))
    return dest, it, {"final_rsq": frsq.value, "final_relrsq": frel.value, "converged": conv.value}


def multicg(src, offsets, resid, parity=EVEN, niter=2000):
    src = np.ascontiguousarray(src, dtype=np.float64)
    n = len(offsets)
    psim = [np.zeros_like(src) for _ in range(n)]
    ptrs = (ctypes.c_void_p * n)(*[p.ctypes.data for p in psim])
    offs = np.ascontiguousarray(offsets, dtype=np.float64)
    frsq = ctypes.c_double(0)
    it = lib().ref_multicg(_p(src), ptrs, _p(offs), n, parity, niter, resid, ctypes.byref(frsq))
    return psim, it, frsq.value


def force_oprod(multi_x, residues, V):
    n = len(multi_x)
    xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
    ptrs = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
    res = np.ascontiguousarray(residues, dtype=np.float64)
    o1 = np.zeros((V, 4, 3, 3, 2))
    o3 = np.zeros((V, 4, 3, 3, 2))
    lib().ref_force_oprod(ptrs, _p(res), n, _p(o1), _p(o3))
    return o1, o3


def hisq_force(multi_x, residues, eps, V, n_naik=0):
    """MILC's HISQ thin-link force on the links of the last hisq_links call
    (n_naik = 0) or hisq_links_naik call (the last n_naik terms in the second
    Naik set); returns the momentum added to zero, (V, 4, 10) packed
    anti_hermitmat."""
    n = len(multi_x)
    xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
    ptrs = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
    res = np.ascontiguousarray(residues, dtype=np.float64)
    mom = np.zeros((V, 4, 10))
    rc = lib().ref_hisq_force_naik(ptrs, _p(res), n - n_naik, n_naik, eps, _p(mom))
    if rc != 0:
        raise RuntimeError("ref_hisq_force: no HISQ links (call hisq_links first)")
    return mom

"""ctypes front end of the FP64 oracle (libks_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module.  All fields use MILC's layout and site order (see ks_oracle.h):
vectors are float64 arrays of shape (V, 3, 2) [re/im], link fields (V, 4, 3, 3, 2),
site index = MILC ``node_index`` (even sites first).
"""
import ctypes
import os
import subprocess

import numpy as np

EVEN, ODD, EVENANDODD = 2, 1, 3

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libks_oracle.so")
_lib = None


class _Geom(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("V", ctypes.c_size_t), ("Vh", ctypes.c_size_t),
                ("coords", ctypes.POINTER(ctypes.c_int)), ("nbr", ctypes.POINTER(ctypes.c_size_t))]


class Qic(ctypes.Structure):
    """Mirror of quark_invert_control's solver fields (include/generic_quark_types.h:167-190)."""
    _fields_ = [("max", ctypes.c_int), ("nrestart", ctypes.c_int), ("parity", ctypes.c_int),
                ("resid", ctypes.c_double), ("relresid", ctypes.c_double),
                ("final_rsq", ctypes.c_double), ("final_relrsq", ctypes.c_double),
                ("size_r", ctypes.c_double), ("size_relr", ctypes.c_double),
                ("final_iters", ctypes.c_int), ("final_restart", ctypes.c_int),
                ("converged", ctypes.c_int)]


def build():
    """Compile libks_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "libks_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        G = ctypes.POINTER(_Geom)
        L.kso_geom_init.argtypes = [G, ctypes.POINTER(ctypes.c_int)]
        L.kso_geom_free.argtypes = [G]
        L.kso_node_index.argtypes = [G, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.kso_node_index.restype = ctypes.c_size_t
        L.kso_gauge_random.argtypes = [G, ctypes.c_uint64, ctypes.c_double, P]
        L.kso_gaussian_vector.argtypes = [G, ctypes.c_uint64, ctypes.c_int, P]
        L.kso_rephase.argtypes = [G, P]
        L.kso_plaquette.argtypes = [G, P]
        L.kso_plaquette.restype = ctypes.c_double
        L.kso_fat_links.argtypes = [G, P, P, P]
        L.kso_long_links.argtypes = [G, ctypes.c_double, P, P]
        L.kso_u3_unitarize.argtypes = [P, P]
        L.kso_hisq_links.argtypes = [G, P, P, P, P]
        L.kso_asqtad_links.argtypes = [G, ctypes.c_double, P, P, P]
        L.kso_dslash.argtypes = [G, P, P, P, P, ctypes.c_int]
        L.kso_congrad.argtypes = [G, P, P, P, P, ctypes.POINTER(Qic), ctypes.c_double]
        L.kso_multicg_offset.argtypes = [G, P, P, P, P, P, ctypes.c_int, ctypes.POINTER(Qic)]
        L.kso_force_oprod.argtypes = [G, P, P, ctypes.c_int, P, P]
        L.kso_fn_force.argtypes = [G, P, P, P, P, ctypes.c_int, ctypes.c_double, P]
        L.kso_hisq_force.argtypes = [G, P, P, P, ctypes.c_int, ctypes.c_double, P]
        L.kso_hisq_force_naik.argtypes = [G, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, P]
        L.kso_hisq_links_naik.argtypes = [G, ctypes.c_double, P, P, P]
        L.kso_u3_adjoint.argtypes = [P, P, P]
        L.kso_set_force_form.argtypes = [ctypes.c_int]
        L.kso_gauge_coeffs.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, P]
        L.kso_gauge_loops.argtypes = [P, P, P]
        L.kso_gauge_action.argtypes = [G, P, P]
        L.kso_gauge_action.restype = ctypes.c_double
        L.kso_gauge_force.argtypes = [G, P, ctypes.c_double, P, P]
        L.kso_mom_action.argtypes = [G, P]
        L.kso_mom_action.restype = ctypes.c_double
        L.kso_update_u.argtypes = [G, ctypes.c_double, P, P]
        L.kso_norm2.argtypes = [G, P, ctypes.c_int]
        L.kso_norm2.restype = ctypes.c_double
        L.kso_mdagm.argtypes = [G, P, P, ctypes.c_double, P, P, P, ctypes.c_int]
        L.kso_prn_seed.argtypes = [G, ctypes.c_int, P]
        L.kso_prn_seed_one.argtypes = [ctypes.c_int, ctypes.c_int, P]
        L.kso_prn_next.argtypes = [P]
        L.kso_prn_next.restype = ctypes.c_uint32
        L.kso_ranmom_prn.argtypes = [G, P, ctypes.c_int, P]
        L.kso_gaussian_prn.argtypes = [G, P, ctypes.c_int, ctypes.c_int, P]
        L.kso_gaussian_prn.restype = ctypes.c_double
        L.kso_reunit_su3.argtypes = [P, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.kso_reunitarize.argtypes = [G, ctypes.c_double, P, ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


class Lattice:
    """Geometry of an nx*ny*nz*nt lattice in MILC site order."""

    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        self._g = _Geom()
        d = (ctypes.c_int * 4)(*self.dims)
        if lib().kso_geom_init(ctypes.byref(self._g), d) != 0:
            raise ValueError("lattice dims must have even nx and even volume")
        self.V = int(self._g.V)
        self.Vh = int(self._g.Vh)
        self.coords = np.ctypeslib.as_array(self._g.coords, shape=(self.V, 4)).copy()

    def __del__(self):
        try:
            lib().kso_geom_free(ctypes.byref(self._g))
        except Exception:
            pass

    @property
    def g(self):
        return ctypes.byref(self._g)

    def node_index(self, x, y, z, t):
        return int(lib().kso_node_index(self.g, x, y, z, t))

    def lex_of_site(self):
        """lexicographic index x + nx(y + ny(z + nz t)) of each MILC site."""
        c = self.coords.astype(np.int64)
        nx, ny, nz, _ = self.dims
        return c[:, 0] + nx * (c[:, 1] + ny * (c[:, 2] + nz * c[:, 3]))

    def parity_slice(self, parity):
        if parity == EVEN:
            return slice(0, self.Vh)
        if parity == ODD:
            return slice(self.Vh, self.V)
        return slice(0, self.V)

    # ---- fields ---------------------------------------------------------
    def new_links(self):
        return np.zeros((self.V, 4, 3, 3, 2))

    def new_vector(self):
        return np.zeros((self.V, 3, 2))

    def gauge_random(self, seed, scale=0.4):
        U = self.new_links()
        lib().kso_gauge_random(self.g, seed, scale, _p(U))
        return U

    def gaussian_vector(self, seed, parity=EVEN):
        v = self.new_vector()
        lib().kso_gaussian_vector(self.g, seed, parity, _p(v))
        return v

    def rephase(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        lib().kso_rephase(self.g, _p(U))
        return U

    def plaquette(self, U):
        return lib().kso_plaquette(self.g, _p(np.ascontiguousarray(U)))

    def fat_links(self, coeff, U):
        fat = self.new_links()
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        lib().kso_fat_links(self.g, _p(c), _p(np.ascontiguousarray(U)), _p(fat))
        return fat

    def long_links(self, naik, U):
        lng = self.new_links()
        lib().kso_long_links(self.g, naik, _p(np.ascontiguousarray(U)), _p(lng))
        return lng

    def hisq_links(self, U_phased, want_W=False):
        fat, lng, W = self.new_links(), self.new_links(), self.new_links()
        nsvd = lib().kso_hisq_links(self.g, _p(np.ascontiguousarray(U_phased)), _p(fat), _p(lng), _p(W))
        return (fat, lng, nsvd, W) if want_W else (fat, lng, nsvd)

    def asqtad_links(self, U_phased, u0):
        fat, lng = self.new_links(), self.new_links()
        lib().kso_asqtad_links(self.g, u0, _p(np.ascontiguousarray(U_phased)), _p(fat), _p(lng))
        return fat, lng

    def dslash(self, fat, lng, src, parity, dst=None):
        if dst is None:
            dst = self.new_vector()
        lib().kso_dslash(self.g, _p(fat), _p(lng), _p(src), _p(dst), parity)
        return dst

    def mdagm(self, fat, lng, offset, src, parity):
        dst, tmp = self.new_vector(), self.new_vector()
        lib().kso_mdagm(self.g, _p(fat), _p(lng), offset, _p(src), _p(dst), _p(tmp), parity)
        return dst

    def norm2(self, v, parity=EVENANDODD):
        return lib().kso_norm2(self.g, _p(v), parity)

    def congrad(self, fat, lng, src, mass, resid, parity=EVEN, max_iters=1000, nrestart=5,
                dest=None, relresid=0.0):
        if dest is None:
            dest = self.new_vector()
        q = Qic(max=max_iters, nrestart=nrestart, parity=parity, resid=resid, relresid=relresid)
        it = lib().kso_congrad(self.g, _p(fat), _p(lng), _p(src), _p(dest), ctypes.byref(q), mass)
        return dest, it, q

    def multicg_offset(self, fat, lng, src, offsets, resid, parity=EVEN, max_iters=1000, nrestart=5):
        n = len(offsets)
        psim = [self.new_vector() for _ in range(n)]
        ptrs = (ctypes.c_void_p * n)(*[p.ctypes.data for p in psim])
        offs = np.ascontiguousarray(offsets, dtype=np.float64)
        qics = (Qic * n)()
        for q in qics:
            q.max, q.nrestart, q.parity, q.resid = max_iters, nrestart, parity, resid
        it = lib().kso_multicg_offset(self.g, _p(fat), _p(lng), _p(src), ptrs, _p(offs), n, qics)
        return psim, it, list(qics)

    def fn_force(self, coeff, U_phased, multi_x, residues, eps=1.0):
        """fermion_force_fn_multi: TA(2 eps U G) per link, full 3x3 matrices (V, 4, 3, 3, 2)."""
        n = len(multi_x)
        ptrs = (ctypes.c_void_p * n)(*[np.ascontiguousarray(x).ctypes.data for x in multi_x])
        res = np.ascontiguousarray(residues, dtype=np.float64)
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        mom = self.new_links()
        lib().kso_fn_force(self.g, _p(c), _p(np.ascontiguousarray(U_phased)), ptrs, _p(res), n, eps, _p(mom))
        return mom

    def gauge_action(self, coeff, U):
        """sum over loops and sites of coeff (3 - Re Tr P) (bare links)"""
        return lib().kso_gauge_action(self.g, _p(np.ascontiguousarray(coeff, dtype=np.float64)),
                                      _p(np.ascontiguousarray(U)))

    def gauge_force(self, coeff, eb3, U, mom):
        """mom (V, 4, 3, 3, 2) full anti-Hermitian += -eb3 TA(U G) (in place)"""
        assert mom.dtype == np.float64 and mom.flags.c_contiguous
        lib().kso_gauge_force(self.g, _p(np.ascontiguousarray(coeff, dtype=np.float64)), eb3,
                              _p(np.ascontiguousarray(U)), _p(mom))
        return mom

    def mom_action(self, mom):
        return lib().kso_mom_action(self.g, _p(np.ascontiguousarray(mom)))

    def update_u(self, eps, U, mom):
        """U <- exp_8(eps P) U in place"""
        assert U.dtype == np.float64 and U.flags.c_contiguous
        lib().kso_update_u(self.g, eps, _p(U), _p(np.ascontiguousarray(mom)))
        return U

    def prn_seed(self, iseed):
        """MILC's per-site generators seeded as initialize_site_prn_from_seed:
        uint64 state (V, 9) in node_index order"""
        st = np.zeros((self.V, 9), dtype=np.uint64)
        lib().kso_prn_seed(self.g, iseed, st.ctypes.data_as(ctypes.c_void_p))
        return st

    def ranmom_prn(self, st, single=True):
        """ranmom() from the site generators st (advanced in place): packed
        anti_hermitmat momenta (V, 4, 10)"""
        assert st.dtype == np.uint64 and st.flags.c_contiguous
        mom = np.zeros((self.V, 4, 10))
        lib().kso_ranmom_prn(self.g, st.ctypes.data_as(ctypes.c_void_p), int(single), _p(mom))
        return mom

    def gaussian_prn(self, st, parity=EVEN, single=True):
        """three complex Gaussians per site of parity from st: (vector, sum |v|^2)"""
        assert st.dtype == np.uint64 and st.flags.c_contiguous
        v = self.new_vector()
        s = lib().kso_gaussian_prn(self.g, st.ctypes.data_as(ctypes.c_void_p), int(single), parity, _p(v))
        return v, s

    def reunitarize(self, U, tol=1e-4):
        """reunit_su3 on every (bare) link in place: (error count, max deviation)"""
        assert U.dtype == np.float64 and U.flags.c_contiguous
        md = ctypes.c_double(0)
        n = lib().kso_reunitarize(self.g, tol, _p(U), ctypes.byref(md))
        return n, md.value

    def hisq_force(self, U_phased, multi_x, residues, eps=1.0, n_naik=0, eps_naik=0.0):
        """HISQ thin-link force TA(2 eps U F_U), full 3x3 matrices (V, 4, 3, 3, 2);
        the last n_naik terms belong to a second Naik set with eps_naik."""
        n = len(multi_x)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
        ptrs = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        res = np.ascontiguousarray(residues, dtype=np.float64)
        mom = self.new_links()
        lib().kso_hisq_force_naik(self.g, _p(np.ascontiguousarray(U_phased)), ptrs, _p(res), n, n_naik, eps_naik,
                                  eps, _p(mom))
        return mom

    def hisq_links_naik(self, U_phased, eps_naik):
        """links of a second Naik set: X_2 + eps_naik X_3 (fat, lng)"""
        fat, lng, _, W = self.hisq_links(U_phased, want_W=True)
        lib().kso_hisq_links_naik(self.g, eps_naik, _p(W), _p(fat), _p(lng))
        return fat, lng

    def force_oprod(self, multi_x, residues):
        n = len(multi_x)
        ptrs = (ctypes.c_void_p * n)(*[np.ascontiguousarray(x).ctypes.data for x in multi_x])
        res = np.ascontiguousarray(residues, dtype=np.float64)
        o1, o3 = self.new_links(), self.new_links()
        lib().kso_force_oprod(self.g, ptrs, _p(res), n, _p(o1), _p(o3))
        return o1, o3


def prn_stream(seed, index, n):
    """n raw 24-bit draws of one MILC generator seeded with (seed, index)"""
    st = np.zeros(9, dtype=np.uint64)
    lib().kso_prn_seed_one(seed, index, st.ctypes.data_as(ctypes.c_void_p))
    return np.array([lib().kso_prn_next(st.ctypes.data_as(ctypes.c_void_p)) for _ in range(n)], dtype=np.uint32)


def unpack_ah(p):
    """packed anti_hermitmat [..., 10] -> full matrices [..., 3, 3, 2]"""
    F = np.zeros(p.shape[:-1] + (3, 3, 2))
    F[..., 0, 1, :] = p[..., 0:2]
    F[..., 0, 2, :] = p[..., 2:4]
    F[..., 1, 2, :] = p[..., 4:6]
    F[..., 1, 0, 0], F[..., 1, 0, 1] = -p[..., 0], p[..., 1]
    F[..., 2, 0, 0], F[..., 2, 0, 1] = -p[..., 2], p[..., 3]
    F[..., 2, 1, 0], F[..., 2, 1, 1] = -p[..., 4], p[..., 5]
    F[..., 0, 0, 1], F[..., 1, 1, 1], F[..., 2, 2, 1] = p[..., 6], p[..., 7], p[..., 8]
    return F


def u3_unitarize(V):
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(3, 3, 2)
    W = np.zeros_like(V)
    svd = lib().kso_u3_unitarize(_p(V), _p(W))
    return W, svd


def gauge_coeffs(u0, nf, hisq=True):
    """loop coefficients (plaquette, rectangle, parallelogram) of the
    Symanzik 1-loop action with HISQ (or asqtad) quark loops"""
    c = np.zeros(3)
    lib().kso_gauge_coeffs(u0, int(nf), int(bool(hisq)), _p(c))
    return c


def gauge_loops():
    """(dirs (n, 6) MILC codes, -1 padded; lengths; shapes) of the action's loops"""
    d = np.zeros(64 * 6, dtype=np.int32)
    ln = np.zeros(64, dtype=np.int32)
    sh = np.zeros(64, dtype=np.int32)
    n = lib().kso_gauge_loops(d.ctypes.data_as(ctypes.c_void_p), ln.ctypes.data_as(ctypes.c_void_p),
                              sh.ctypes.data_as(ctypes.c_void_p))
    return d.reshape(64, 6)[:n], ln[:n], sh[:n]


def set_force_form(form):
    """0: force adjoints through the path tables, 1: staple (reverse-mode) form"""
    lib().kso_set_force_form(int(form))


def u3_adjoint(V, FW):
    """force through the U(3) projection at one link: F_V from F_W ((3,3,2) arrays)"""
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(3, 3, 2)
    FW = np.ascontiguousarray(FW, dtype=np.float64).reshape(3, 3, 2)
    FV = np.zeros_like(V)
    lib().kso_u3_adjoint(_p(V), _p(FW), _p(FV))
    return FV


def to_complex(a):
    """(..., 2) float array -> complex array."""
    return a[..., 0] + 1j * a[..., 1]


def from_complex(c):
    return np.stack([c.real, c.imag], axis=-1).astype(np.float64)


# HISQ level-2 free-field coefficients: fat = 9/8, Naik = -1/24 (see kso_hisq_links)
HISQ_FREE_C1 = 9.0 / 8.0


def asqtad_coeffs(u0):
    """act_path_coeff {one_link, naik, 3-staple, 5-staple, 7-staple, Lepage} of
    generic_ks/imp_actions/asqtad_action.h with tadpole factors u0^-(length-1)."""
    u2 = u0 * u0
    return [5.0 / 8.0, -1.0 / 24.0 / u2, -1.0 / 16.0 / u2, 1.0 / 64.0 / u2 ** 2, -1.0 / 384.0 / u2 ** 3,
            -1.0 / 16.0 / u2 ** 2]
HISQ_FREE_C3 = -1.0 / 24.0

"""Host-side mirror of MILC's generic_ks GPU interface over the C-ABI.

Every call goes through ``libmilc_hip.so`` (include/milc_hip.h).  The Python
names follow the MILC functions they stand in for, with the same argument
meaning and the same error behaviour (non-convergence is reported in the
returned control record, invalid arguments raise):

    ks_congrad_parity      generic_ks/d_congrad5_fn_gpu.c:35  (-> milc_hip_invert)
    ks_multicg_offset_field generic_ks/ks_multicg_offset_gpu.c:38 (-> milc_hip_multishift_invert)
    dslash_fn_field        generic_ks/dslash_fn.c:306  (-> milc_hip_dslash)
    compute_oprod          generic_ks/fermion_force_asqtad_gpu.c:60 (-> milc_hip_compute_oprod)
    load_hisq_links        generic_ks/fermion_links_hisq_load_milc.c:546-702 (-> milc_hip_load_hisq_links_naik;
                           eps_naik != 0: the Naik set of load_hisq_fn_links, :574-640)

Host fields are numpy arrays in MILC's own layout and site order: vectors
(V, 3, 2), links (V, 4, 3, 3, 2), even sites first; float64 (MILC_PRECISION
2) throughout, and float32 (MILC_PRECISION 1) for the links and fields of the
Dslash, CG and multi-shift calls.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

EVEN, ODD, EVENANDODD = 2, 1, 3  # include/macros.h:17-19

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmilc_hip.so")
_lib = None
_active = None  # the Lattice whose geometry the library context currently holds


class MilcHipError(RuntimeError):
    pass


class InvArgs(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("evenodd", ctypes.c_int), ("mixed_precision", ctypes.c_int),
                ("naik_epsilon", ctypes.c_double), ("tadpole", ctypes.c_double),
                ("restart_iters", ctypes.c_int)]


EXPORTS = [
    "milc_hip_init", "milc_hip_finalize", "milc_hip_version", "milc_hip_invert",
    "milc_hip_multishift_invert", "milc_hip_dslash", "milc_hip_compute_oprod",
    "milc_hip_load_hisq_links", "milc_hip_load_hisq_links_naik", "milc_hip_dev_gauge_hisq", "milc_hip_dev_source",
    "milc_hip_dev_invert", "milc_hip_dev_read", "milc_hip_dev_write", "milc_hip_dev_read_links",
    "milc_hip_dev_dslash_bench", "milc_hip_dev_copy_bench", "milc_hip_dev_diff", "milc_hip_dev_cg_iteration_bench",
    "milc_hip_dev_rhmc_action", "milc_hip_dev_force_level0", "milc_hip_dev_load_links",
    "milc_hip_asqtad_force", "milc_hip_fermion_force_fn_multi",
    "milc_hip_invert_block", "milc_hip_multishift_invert_block", "milc_hip_dev_block_tune",
    "milc_hip_dev_block_dslash_bench", "milc_hip_dev_block_dslash", "milc_hip_dev_multishift_block",
    "milc_hip_dev_block_describe",
    "milc_hip_dev_multishift", "milc_hip_dev_hmc_step", "milc_hip_hisq_force", "milc_hip_eo_fermion_force_multi",
    "milc_hip_set_force_form", "milc_hip_dev_rhmc_trajectory", "milc_hip_gauge_force", "milc_hip_gauge_action",
    "milc_hip_update_u", "milc_hip_mom_action", "milc_hip_dev_load_thin_links", "milc_hip_dev_read_thin_links",
    "milc_hip_dev_dslash", "milc_hip_set_link_compression", "milc_hip_link_compression", "milc_hip_dev_hisq_force",
    "milc_hip_reunitarize", "milc_hip_dev_seed_rng", "milc_hip_dev_ranmom", "milc_hip_dev_gaussian_source",
    "milc_hip_debug_fail_trajectory",
]


class Rational(ctypes.Structure):
    """MILC's params_ratfunc: R(x) = res[0] + sum_k res[k] / (x + pole[k])"""
    _fields_ = [("order", ctypes.c_int), ("res", ctypes.c_double * 33), ("pole", ctypes.c_double * 33)]

    @classmethod
    def make(cls, norm, residues, poles):
        r = cls()
        r.order = len(residues)
        r.res[0] = norm
        for k, (a, b) in enumerate(zip(residues, poles)):
            r.res[k + 1] = a
            r.pole[k + 1] = b
        return r


def load_rhmc_params(path):
    """MILC's rational-function parameter file (ks_imp_rhmc/load_rhmc_params.c):
    one dict per pseudofermion, {"MD"|"GR"|"FA": Rational, "naik_epsilon": float},
    res[0] = constant; naik_term_epsilon precedes each pseudofermion's rationals"""
    out, cur, raw, eps = [], None, {}, []
    pending = 0.0
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "naik_term_epsilon":
            pending = float(t[1])
        for kind in ("MD", "GR", "FA"):
            if t[0] == "order_" + kind:
                if kind == "MD":
                    raw = {}
                    out.append(raw)
                    eps.append(pending)
                    pending = 0.0
                raw[kind] = {"order": int(t[1]), "res": [], "pole": []}
            elif t[0] == "res_" + kind:
                raw[kind]["res"].append(float(t[1]))
            elif t[0] == "pole_" + kind:
                raw[kind]["pole"].append(float(t[1]))
    for raw in out:
        for kind, v in raw.items():
            if len(v["res"]) != v["order"] + 1 or len(v["pole"]) != v["order"] + 1:
                raise ValueError(f"{path}: {kind} has order {v['order']} but {len(v['res'])} residues")
            raw[kind] = Rational.make(v["res"][0], v["res"][1:], v["pole"][1:])
    for raw, e in zip(out, eps):
        raw["naik_epsilon"] = e
    return out


def gauge_loop_coeffs(u0, nf, hisq=True):
    """(plaquette, 1x2 rectangle, 1x1x1 parallelogram) coefficients of the
    Symanzik 1-loop gauge action with HISQ (nf dynamical flavours) or asqtad
    quark loops (generic/imp_actions/symanzik_1loop_hisq_action.h,
    symanzik_1loop_action.h)"""
    import math
    lg = math.log(u0)
    if hisq:
        return [1.0, -1.0 / (20.0 * u0 * u0) * (1.0 - (0.6264 - 1.1746 * nf) * lg),
                1.0 / (u0 * u0) * (0.0433 - 0.0156 * nf) * lg]
    return [1.0, -1.0 / (20.0 * u0 * u0) * (1.0 - 0.6264 * lg), 1.0 / (u0 * u0) * 0.04335 * lg]


MAX_PSEUDO = 8


class Pseudofermion(ctypes.Structure):
    """one pseudofermion: mass (4 m^2 added to the poles; 0 for MILC's files,
    whose poles hold it), MD / GR / FA rationals, solver targets"""
    _fields_ = [("mass", ctypes.c_double), ("md", Rational), ("gr", Rational), ("fa", Rational),
                ("resid_md", ctypes.c_double), ("resid_gr", ctypes.c_double), ("resid_fa", ctypes.c_double),
                ("naik_epsilon", ctypes.c_double)]


class RhmcParams(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("integrator", ctypes.c_int), ("epsilon", ctypes.c_double),
                ("beta", ctypes.c_double), ("loop_coeff", ctypes.c_double * 3), ("action", ctypes.c_int),
                ("u0", ctypes.c_double), ("n_pseudo", ctypes.c_int),
                ("pf", Pseudofermion * MAX_PSEUDO), ("mixed_md", ctypes.c_int), ("mixed_action", ctypes.c_int),
                ("action_from_md", ctypes.c_int),
                ("accept", ctypes.c_int), ("site_rng", ctypes.c_int), ("seed", ctypes.c_ulonglong)]


# MILC's int_alg_t (ks_imp_rhmc/ks_imp_includes.h:58-60)
INT_LEAPFROG, INT_OMELYAN, INT_2G1F, INT_3G1F, INT_5G1F, INT_6G1F = 0, 1, 4, 5, 6, 7

TRAJECTORY_STATS = ["g0", "h0", "f0", "g1", "h1", "f1", "dH", "accepted", "seconds", "md_iters", "fermion_forces",
                    "fermion_force_seconds", "gauge_force_seconds", "link_seconds", "action_seconds", "gr_iters",
                    "gauge_forces", "reunit_max_deviation", "reunit_errors", "link_rebuilds"]


def pseudofermions_from_file(path, resid_md, resid_gr, resid_fa, mass=0.0):
    """Pseudofermion list from MILC's rational file; resid_* are per-pseudofermion
    lists (or one value for all), MILC's cgresid_md_fa_gr lines"""
    rats = load_rhmc_params(path)
    n = len(rats)

    def per(v):
        return list(v) if hasattr(v, "__len__") else [v] * n
    rm, rg, rf, ms = per(resid_md), per(resid_gr), per(resid_fa), per(mass)
    return [Pseudofermion(mass=ms[i], md=r["MD"], gr=r["GR"], fa=r["FA"], resid_md=rm[i], resid_gr=rg[i],
                          resid_fa=rf[i], naik_epsilon=r["naik_epsilon"]) for i, r in enumerate(rats)]


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MilcHipError(f"{path} is missing: build it with __graft_entry__.build() or "
                           "make -C milc_qcd_amd/csrc")
    L = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.milc_hip_init.argtypes = [ctypes.POINTER(ctypes.c_int), I]
    L.milc_hip_version.restype = ctypes.c_char_p
    L.milc_hip_invert.argtypes = [I, I, D, InvArgs, D, D, P, P, P, P, ctypes.POINTER(D), ctypes.POINTER(D),
                                  ctypes.POINTER(I)]
    L.milc_hip_multishift_invert.argtypes = [I, I, I, ctypes.POINTER(D), InvArgs, ctypes.POINTER(D),
                                             ctypes.POINTER(D), P, P, P, ctypes.POINTER(P), ctypes.POINTER(D),
                                             ctypes.POINTER(D), ctypes.POINTER(I)]
    L.milc_hip_dslash.argtypes = [I, I, InvArgs, P, P, P, P, ctypes.POINTER(I)]
    L.milc_hip_compute_oprod.argtypes = [I, I, ctypes.POINTER(ctypes.POINTER(D)), ctypes.POINTER(P),
                                         ctypes.POINTER(P)]
    L.milc_hip_load_hisq_links.argtypes = [I, P, P, P, P, ctypes.POINTER(I)]
    L.milc_hip_load_hisq_links_naik.argtypes = [I, P, ctypes.c_double, P, P, P, ctypes.POINTER(I)]
    L.milc_hip_dev_gauge_hisq.argtypes = [ctypes.c_uint64, D, ctypes.POINTER(I)]
    L.milc_hip_dev_source.argtypes = [ctypes.c_uint64, I, I]
    L.milc_hip_dev_invert.argtypes = [I, I, D, I, D, I, I, ctypes.POINTER(D)]
    L.milc_hip_dev_invert_noalloc.argtypes = [I, I, D, I, D, I, I, ctypes.POINTER(D)]
    L.milc_hip_dev_read.argtypes = [I, I, P]
    L.milc_hip_dev_write.argtypes = [I, I, P]
    L.milc_hip_dev_read_links.argtypes = [P, P]
    L.milc_hip_dev_dslash_bench.argtypes = [I, I, I, ctypes.POINTER(D)]
    L.milc_hip_dev_cg_iteration_bench.argtypes = [I, D, I, I, ctypes.POINTER(D)]
    L.milc_hip_dev_copy_bench.argtypes = [ctypes.c_size_t, I, ctypes.POINTER(D)]
    L.milc_hip_dev_diff.argtypes = [I, I, ctypes.POINTER(D)]
    L.milc_hip_dev_rhmc_action.argtypes = [I, I, ctypes.POINTER(D), ctypes.POINTER(D), D, I, ctypes.POINTER(D)]
    L.milc_hip_dev_force_level0.argtypes = [I, I, ctypes.POINTER(D), ctypes.POINTER(D), D, I, ctypes.POINTER(D), P, P]
    L.milc_hip_dev_load_links.argtypes = [P, P]
    L.milc_hip_asqtad_force.argtypes = [I, ctypes.POINTER(D), ctypes.POINTER(P), ctypes.POINTER(P), P, P]
    L.milc_hip_fermion_force_fn_multi.argtypes = [I, D, ctypes.POINTER(D), ctypes.POINTER(P), I,
                                                  ctypes.POINTER(D), P, P]
    PI = ctypes.POINTER(I)
    PD = ctypes.POINTER(D)
    L.milc_hip_invert_block.argtypes = [I, I, D, InvArgs, D, P, P, ctypes.POINTER(P), ctypes.POINTER(P), PD, PI]
    L.milc_hip_multishift_invert_block.argtypes = [I, I, I, PD, InvArgs, PD, D, P, P, ctypes.POINTER(P),
                                                   ctypes.POINTER(P), PD, PI]
    L.milc_hip_dev_block_tune.argtypes = [I, I, PI, PD, ctypes.c_char_p, I]
    L.milc_hip_dev_block_dslash_bench.argtypes = [I, I, I, I, PD]
    L.milc_hip_dev_block_describe.argtypes = [I, I, I, ctypes.c_char_p, I]
    L.milc_hip_dev_block_dslash.argtypes = [I, I, I, I, PI, I, PI]
    L.milc_hip_dev_multishift_block.argtypes = [I, PI, I, PD, D, D, I, I, I, I, PD]
    L.milc_hip_dev_multishift.argtypes = [I, I, I, PD, D, I, I, PD]
    L.milc_hip_dev_hmc_step.argtypes = [I, I, PD, PD, D, I, I, I, PD]
    L.milc_hip_hisq_force.argtypes = [I, I, I, D, ctypes.POINTER(PD), ctypes.POINTER(P), PD, PD, P, P, P, P]
    L.milc_hip_eo_fermion_force_multi.argtypes = [I, D, PD, ctypes.POINTER(P), I, P, P]
    L.milc_hip_set_force_form.argtypes = [I]
    L.milc_hip_dev_rhmc_trajectory.argtypes = [ctypes.POINTER(RhmcParams), ctypes.POINTER(D)]
    L.milc_hip_debug_fail_trajectory.argtypes = [I]
    L.milc_hip_dev_seed_rng.argtypes = [I, I]
    L.milc_hip_dev_ranmom.argtypes = [P]
    L.milc_hip_dev_gaussian_source.argtypes = [I, I, ctypes.POINTER(D)]
    L.milc_hip_reunitarize.argtypes = [I, D, P, I, ctypes.POINTER(D), ctypes.POINTER(I)]
    L.milc_hip_gauge_force.argtypes = [I, I, ctypes.POINTER(D), D, P, I, P]
    L.milc_hip_gauge_action.argtypes = [I, I, ctypes.POINTER(D), P, I, ctypes.POINTER(D)]
    L.milc_hip_update_u.argtypes = [I, D, P, P, I]
    L.milc_hip_mom_action.argtypes = [I, P, ctypes.POINTER(D)]
    L.milc_hip_dev_load_thin_links.argtypes = [I, P, I, ctypes.POINTER(I)]
    L.milc_hip_dev_read_thin_links.argtypes = [P, I]
    L.milc_hip_dev_dslash.argtypes = [I, I, I, I]
    L.milc_hip_set_link_compression.argtypes = [I]
    L.milc_hip_link_compression.argtypes = [I]
    L.milc_hip_dev_hisq_force.argtypes = [I, I, P, ctypes.c_double, P, P]
    _lib = L
    return L


def _real(a):
    """MILC_PRECISION of a host field: 1 for float32 arrays, 2 for float64."""
    if a.dtype == np.float32:
        return 1
    if a.dtype == np.float64:
        return 2
    raise MilcHipError("fields must be float32 (MILC_PRECISION 1) or float64 (MILC_PRECISION 2) arrays")


def _pr(a):
    """Pointer to a C-contiguous float32 or float64 host field."""
    _real(a)
    if not a.flags.c_contiguous:
        raise MilcHipError("fields must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def _p(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise MilcHipError("fields must be C-contiguous float64 arrays")
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc < 0:
        raise MilcHipError(f"{what} failed with status {rc}")
    return rc


def _evenodd(parity):
    if parity == EVEN:
        return 0
    if parity == ODD:
        return 1
    # the GPU glue rejects EVENANDODD (d_congrad5_fn_gpu.c:95-102)
    raise MilcHipError("parity must be EVEN or ODD")


@dataclass
class QuarkInvertControl:
    """Solver fields of quark_invert_control (include/generic_quark_types.h:167-190)."""
    max: int = 1000
    nrestart: int = 5
    parity: int = EVEN
    resid: float = 1e-8
    relresid: float = 0.0
    prec: int = 2
    mixed: int = 0  # 0 plain, 1 FP32 sloppy, 2 16-bit sloppy (with FP64 reliable updates)
    final_rsq: float = 0.0
    final_relrsq: float = 0.0
    size_r: float = 0.0
    size_relr: float = 0.0
    converged: int = 0
    final_iters: int = 0
    final_restart: int = 0


@dataclass
class KsParam:
    """ks_param (include/generic_quark_types.h:131-139); offset includes 4 m^2."""
    offset: float
    residue: float = 1.0
    mass: float = 0.0
    naik_term_epsilon: float = 0.0


class Lattice:
    """One nx*ny*nz*nt lattice on the device (milc_hip_init)."""

    def __init__(self, dims, device=0):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.Vh = self.V // 2
        self.device = device
        self.lib = load_library()
        self._fresh = True
        self._activate()

    def _activate(self):
        """The library holds one device context (one lattice) per process, as
        MILC's GPU glue does; re-initialise it when another Lattice was used
        since (device-resident fields and links of the other one are dropped)."""
        global _active
        if _active is self:
            return
        d = (ctypes.c_int * 4)(*self.dims)
        _check(self.lib.milc_hip_init(d, self.device), "milc_hip_init")
        _active = self
        self._fresh = True

    def close(self):
        global _active
        if _active is self:
            self.lib.milc_hip_finalize()
            _active = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def parity_slice(self, parity):
        return slice(0, self.Vh) if parity == EVEN else slice(self.Vh, self.V)

    # ---------------------------------------------------------------- links
    def set_links(self, fat, lng):
        """Hold the (fat, long) pair as the current fn links; the next call
        uploads them (MILC's refresh signal, d_congrad5_fn_gpu.c:120-126).
        float32 links are MILC_PRECISION 1 arrays: the operator and solver
        calls then take float32 fields."""
        fat, lng = np.asarray(fat), np.asarray(lng)
        dt = np.float32 if fat.dtype == np.float32 and lng.dtype == np.float32 else np.float64
        self.fat = np.ascontiguousarray(fat, dtype=dt)
        self.lng = np.ascontiguousarray(lng, dtype=dt)
        self._fresh = True

    def _field(self, a):
        """A host field in the precision of the current links."""
        return np.ascontiguousarray(a, dtype=self.fat.dtype)

    def _refresh_flag(self):
        n = ctypes.c_int(-1 if getattr(self, "_fresh", True) else 0)
        self._fresh = False
        return n

    def load_hisq_links(self, U_phased, eps_naik=0.0):
        """HISQ U(3) fat and Naik links from thin links with phases in; with
        eps_naik != 0 the links of that Naik set (X_2 + eps_naik X_3)."""
        self._activate()
        U = np.ascontiguousarray(U_phased, dtype=np.float64)
        fat, lng, W = np.zeros_like(U), np.zeros_like(U), np.zeros_like(U)
        nsvd = ctypes.c_int(0)
        _check(self.lib.milc_hip_load_hisq_links_naik(2, _p(U), ctypes.c_double(eps_naik), _p(fat), _p(lng), _p(W),
                                                      ctypes.byref(nsvd)), "milc_hip_load_hisq_links_naik")
        return fat, lng, W, nsvd.value

    # ---------------------------------------------------------------- operators
    def dslash_fn_field(self, src, parity, prec=2):
        """dest = D src on `parity` (src read on the other parity)."""
        self._activate()
        src = self._field(src)
        dst = np.zeros_like(src)
        args = InvArgs(0, _evenodd(parity), 0, 0.0, 1.0, 0)
        _check(self.lib.milc_hip_dslash(_real(src), prec, args, _pr(self.fat), _pr(self.lng), _pr(src), _pr(dst),
                                        ctypes.byref(self._refresh_flag())), "milc_hip_dslash")
        return dst

    def ks_congrad_parity(self, src, dest, qic: QuarkInvertControl, mass):
        """Solve (M^dag M) dest = src on qic.parity; dest is the initial guess
        (ks_congrad_parity_gpu, d_congrad5_fn_gpu.c:35-166).  Fields are in the
        precision of the links (float32 = MILC_PRECISION 1).  qic.relresid > 0
        adds MILC's Fermilab relative residual test (d_congrad5_fn_milc.c:92):
        final_relrsq is then my_relative_residue of the solution."""
        self._activate()
        src = self._field(src)
        if dest.dtype != self.fat.dtype or not dest.flags.c_contiguous:
            raise MilcHipError(f"dest must be a C-contiguous {self.fat.dtype} array")
        args = InvArgs(qic.max * qic.nrestart, _evenodd(qic.parity), qic.mixed, 0.0, 1.0, qic.max)
        res, rel = ctypes.c_double(0), ctypes.c_double(0)
        n = self._refresh_flag()
        it = _check(self.lib.milc_hip_invert(_real(src), qic.prec, mass, args, qic.resid, qic.relresid,
                                             _pr(self.fat), _pr(self.lng), _pr(src), _pr(dest), ctypes.byref(res),
                                             ctypes.byref(rel), ctypes.byref(n)), "milc_hip_invert")
        qic.final_rsq = res.value ** 2
        qic.final_relrsq = rel.value ** 2
        qic.final_iters = it
        # the glue's convergence flag (d_congrad5_fn_gpu.c:156)
        qic.converged = int(res.value < qic.resid)
        return it

    def ks_congrad_field(self, src, dest, qic: QuarkInvertControl, mass):
        """EVEN / ODD / EVENANDODD dispatch of generic_ks/d_congrad5_fn.c:16-36:
        EVENANDODD solves the even, then the odd sites; qic.parity is restored."""
        parity = qic.parity
        iters = 0
        if parity in (EVEN, EVENANDODD):
            qic.parity = EVEN
            iters += self.ks_congrad_parity(src, dest, qic, mass)
        if parity in (ODD, EVENANDODD):
            qic.parity = ODD
            iters += self.ks_congrad_parity(src, dest, qic, mass)
        qic.parity = parity
        return iters

    def ks_congrad_block_parity(self, srcs, dests, qic: QuarkInvertControl, mass):
        """ks_congrad_block_parity_gpu (d_congrad5_fn_gpu.c:174-176) through the
        batched solver: dests[s] are initial guesses, overwritten on qic.parity.
        Returns (iterations, true residual per source)."""
        self._activate()
        n = len(srcs)
        srcs = [np.ascontiguousarray(x, dtype=np.float64) for x in srcs]
        for d in dests:
            if d.dtype != np.float64 or not d.flags.c_contiguous:
                raise MilcHipError("dests must be C-contiguous float64 arrays")
        sp = (ctypes.c_void_p * n)(*[x.ctypes.data for x in srcs])
        dp = (ctypes.c_void_p * n)(*[d.ctypes.data for d in dests])
        fres = (ctypes.c_double * n)()
        args = InvArgs(qic.max * qic.nrestart, _evenodd(qic.parity), 1, 0.0, 1.0, qic.max)
        it = _check(self.lib.milc_hip_invert_block(2, n, mass, args, qic.resid, _p(self.fat), _p(self.lng), sp, dp,
                                                   fres, ctypes.byref(self._refresh_flag())),
                    "milc_hip_invert_block")
        qic.final_iters = it
        qic.final_rsq = max(fres) ** 2
        qic.converged = int(max(fres) < qic.resid)
        return it, list(fres)

    def ks_multicg_offset_block(self, srcs, offsets, parity, resid, refine_resid, max_iter=20000):
        """Multi-shift for several sources at once; returns (psim[s][j], true
        residual [s][j], iterations)."""
        self._activate()
        ns, n = len(srcs), len(offsets)
        srcs = [np.ascontiguousarray(x, dtype=np.float64) for x in srcs]
        psim = [[np.zeros_like(srcs[0]) for _ in range(n)] for _ in range(ns)]
        sp = (ctypes.c_void_p * ns)(*[x.ctypes.data for x in srcs])
        xp = (ctypes.c_void_p * (ns * n))(*[psim[s][j].ctypes.data for s in range(ns) for j in range(n)])
        offs = (ctypes.c_double * n)(*offsets)
        tr = (ctypes.c_double * 1)(resid)
        fres = (ctypes.c_double * (ns * n))()
        args = InvArgs(max_iter, _evenodd(parity), 1, 0.0, 1.0, 0)
        it = _check(self.lib.milc_hip_multishift_invert_block(2, ns, n, offs, args, tr, refine_resid, _p(self.fat),
                                                              _p(self.lng), sp, xp, fres,
                                                              ctypes.byref(self._refresh_flag())),
                    "milc_hip_multishift_invert_block")
        res = [[fres[s * n + j] for j in range(n)] for s in range(ns)]
        return psim, res, it

    def ks_multicg_offset_field(self, src, ksp, qic, mixed=0, prec=2):
        """Multi-shift CG; returns the list of solutions psim[j]."""
        self._activate()
        src = self._field(src)
        n = len(ksp)
        psim = [np.zeros_like(src) for _ in range(n)]
        offs = (ctypes.c_double * n)(*[k.offset for k in ksp])
        resid = (ctypes.c_double * n)(*[q.resid for q in qic])
        relr = (ctypes.c_double * n)(*[q.relresid for q in qic])
        fres = (ctypes.c_double * n)()
        frel = (ctypes.c_double * n)()
        ptrs = (ctypes.c_void_p * n)(*[p.ctypes.data for p in psim])
        args = InvArgs(qic[0].max * qic[0].nrestart, _evenodd(qic[0].parity), mixed, 0.0, 1.0, 0)
        it = ctypes.c_int(-1 if getattr(self, "_fresh", True) else 0)
        self._fresh = False
        _check(self.lib.milc_hip_multishift_invert(_real(src), prec, n, offs, args, resid, relr, _pr(self.fat),
                                                   _pr(self.lng), _pr(src), ptrs, fres, frel, ctypes.byref(it)),
               "milc_hip_multishift_invert")
        for j, q in enumerate(qic):
            q.final_rsq = fres[j] ** 2
            q.final_iters = it.value
            q.converged = int(fres[j] <= max(q.resid, 1e-14) * 1.0000001) if mixed else 1
        return psim

    def compute_oprod(self, multi_x, coeff):
        """O1, O3 (each (V, 4, 3, 3, 2)) of the level-0 force; coeff[t] = (c1, c3)."""
        self._activate()
        n = len(multi_x)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
        rows = [(ctypes.c_double * 2)(*c) for c in coeff]
        cptr = (ctypes.POINTER(ctypes.c_double) * n)(*[ctypes.cast(r, ctypes.POINTER(ctypes.c_double))
                                                       for r in rows])
        qptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        o1 = [np.zeros((self.V, 3, 3, 2)) for _ in range(4)]
        o3 = [np.zeros((self.V, 3, 3, 2)) for _ in range(4)]
        d1 = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in o1])
        d3 = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in o3])
        op = (ctypes.c_void_p * 2)(ctypes.cast(d1, ctypes.c_void_p), ctypes.cast(d3, ctypes.c_void_p))
        _check(self.lib.milc_hip_compute_oprod(2, n, cptr, qptr, op), "milc_hip_compute_oprod")
        return np.stack(o1, axis=1), np.stack(o3, axis=1)

    def fermion_force_fn_multi(self, eps, residues, multi_x, act_path_coeff, links, mom):
        """fermion_force_fn_multi (generic_ks/fermion_force_fn_multi.c:81-313):
        mom (V, 4, 10) packed anti_hermitmat, updated in place by TA(2 eps U G);
        links = thin links (V, 4, 3, 3, 2) with the KS phases in; multi_x[t] (V, 3, 2)."""
        self._activate()
        n = len(multi_x)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
        qptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        res = (ctypes.c_double * n)(*residues)
        c = (ctypes.c_double * 6)(*act_path_coeff)
        U = np.ascontiguousarray(links, dtype=np.float64)
        _check(self.lib.milc_hip_fermion_force_fn_multi(2, eps, res, qptr, n, c, _p(U), _p(mom)),
               "milc_hip_fermion_force_fn_multi")
        return mom

    def set_link_compression(self, on):
        """1 (default): 16-bit long links in the two-row form when scaled-unitary"""
        self._activate()
        _check(self.lib.milc_hip_set_link_compression(int(on)), "milc_hip_set_link_compression")

    def link_compression(self, prec):
        """long-link format of the device links of precision prec (1: two-row form)"""
        self._activate()
        return _check(self.lib.milc_hip_link_compression(prec), "milc_hip_link_compression")

    def gauge_force(self, loop_coeff, eb3, links, mom, phases_in=False):
        """mom (V, 4, 10) packed += -eb3 TA(U G) (qudaGaugeForcePhased / imp_gauge_force)"""
        self._activate()
        c = (ctypes.c_double * 3)(*loop_coeff)
        U = np.ascontiguousarray(links, dtype=np.float64)
        assert mom.dtype == np.float64 and mom.flags.c_contiguous
        _check(self.lib.milc_hip_gauge_force(2, 3, c, eb3, _p(U), int(bool(phases_in)), _p(mom)),
               "milc_hip_gauge_force")
        return mom

    def gauge_action(self, loop_coeff, links, phases_in=False):
        self._activate()
        c = (ctypes.c_double * 3)(*loop_coeff)
        a = ctypes.c_double(0)
        _check(self.lib.milc_hip_gauge_action(2, 3, c, _p(np.ascontiguousarray(links, dtype=np.float64)),
                                              int(bool(phases_in)), ctypes.byref(a)), "milc_hip_gauge_action")
        return a.value

    def update_u(self, eps, links, mom, phases_in=False):
        """links (V, 4, 3, 3, 2) FP64 updated in place"""
        self._activate()
        assert links.dtype == np.float64 and links.flags.c_contiguous
        _check(self.lib.milc_hip_update_u(2, eps, _p(links), _p(np.ascontiguousarray(mom, dtype=np.float64)),
                                          int(bool(phases_in))), "milc_hip_update_u")
        return links

    def mom_action(self, mom):
        self._activate()
        a = ctypes.c_double(0)
        _check(self.lib.milc_hip_mom_action(2, _p(np.ascontiguousarray(mom, dtype=np.float64)), ctypes.byref(a)),
               "milc_hip_mom_action")
        return a.value

    def dev_rhmc_trajectory(self, steps, epsilon, beta, loop_coeff, pseudofermions, integrator=INT_OMELYAN,
                            mixed_md=1, action_from_md=False, accept=1, seed=1, site_rng=False, mixed_action=0,
                            asqtad_u0=None):
        """one RHMC trajectory on the device links (update(), ks_imp_rhmc/update_rhmc.c);
        pseudofermions: list of Pseudofermion; asqtad_u0: the asqtad quark action with
        that tadpole factor instead of HISQ"""
        self._activate()
        if not 1 <= len(pseudofermions) <= MAX_PSEUDO:
            raise MilcHipError(f"1..{MAX_PSEUDO} pseudofermions")
        p = RhmcParams(steps=steps, integrator=integrator, epsilon=epsilon, beta=beta, n_pseudo=len(pseudofermions),
                       action=0 if asqtad_u0 is None else 1, u0=0.0 if asqtad_u0 is None else asqtad_u0,
                       mixed_md=int(mixed_md), mixed_action=int(mixed_action),
                       action_from_md=int(bool(action_from_md)), accept=accept,
                       site_rng=int(bool(site_rng)), seed=seed)
        for k in range(3):
            p.loop_coeff[k] = loop_coeff[k]
        for i, q in enumerate(pseudofermions):
            p.pf[i] = q
        st = (ctypes.c_double * 32)()
        _check(self.lib.milc_hip_dev_rhmc_trajectory(ctypes.byref(p), st), "milc_hip_dev_rhmc_trajectory")
        out = dict(zip(TRAJECTORY_STATS, list(st)))
        out["grsource"] = list(st)[20:20 + len(pseudofermions)]
        out["action_iters"] = st[28]
        out["setup_seconds"], out["heatbath_seconds"], out["md_solve_seconds"] = st[29], st[30], st[31]
        return out

    def debug_fail_trajectory(self, stage):
        """the next trajectory fails with -90 at stage 1 (heatbath), 2 (start
        actions) or 3 (molecular dynamics): a test hook of the error contract"""
        self._activate()
        _check(self.lib.milc_hip_debug_fail_trajectory(stage), "milc_hip_debug_fail_trajectory")

    def dev_seed_rng(self, iseed, precision=1):
        """MILC's per-site generators (initialize_site_prn_from_seed) and node
        generator; precision 1 rounds the draws as MILC's single-precision build"""
        self._activate()
        _check(self.lib.milc_hip_dev_seed_rng(iseed, precision), "milc_hip_dev_seed_rng")

    def dev_ranmom(self):
        """ranmom() from the site generators: packed momenta (V, 4, 10)"""
        self._activate()
        mom = np.zeros((self.V, 4, 10))
        _check(self.lib.milc_hip_dev_ranmom(_p(mom)), "milc_hip_dev_ranmom")
        return mom

    def dev_gaussian_source(self, parity, slot):
        """three complex Gaussians per site of parity from the site generators
        into device slot; returns MILC's GRSOURCE sum"""
        self._activate()
        s = ctypes.c_double(0)
        _check(self.lib.milc_hip_dev_gaussian_source(parity, slot, ctypes.byref(s)), "milc_hip_dev_gaussian_source")
        return s.value

    def reunitarize(self, links, tolerance=1e-4, phases_in=True):
        """reunitarize_ks on host links (V, 4, 3, 3, 2) in place: (errors, max deviation)"""
        self._activate()
        md, ne = ctypes.c_double(0), ctypes.c_int(0)
        _check(self.lib.milc_hip_reunitarize(2, tolerance, _p(links), int(bool(phases_in)), ctypes.byref(md),
                                             ctypes.byref(ne)), "milc_hip_reunitarize")
        return ne.value, md.value

    def set_force_form(self, form):
        """0: path-table gather; 1 (default): staple form where exact for any
        links, path table otherwise; 2: staple form everywhere (unitary links)"""
        self._activate()
        _check(self.lib.milc_hip_set_force_form(form), "milc_hip_set_force_form")

    def eo_fermion_force_multi(self, eps, residues, multi_x, links, mom):
        """HISQ eo_fermion_force_multi (fermion_force_hisq_multi.c): mom (V, 4, 10)
        += TA(2 eps U F_U); links = thin links with phases in."""
        self._activate()
        n = len(multi_x)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
        qptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        res = (ctypes.c_double * n)(*residues)
        U = np.ascontiguousarray(links, dtype=np.float64)
        _check(self.lib.milc_hip_eo_fermion_force_multi(2, eps, res, qptr, n, _p(U), _p(mom)),
               "milc_hip_eo_fermion_force_multi")
        return mom

    def hisq_force(self, eps, coeff, multi_x, level2_coeff, fat7_coeff, W, V, U, mom, num_naik_terms=0):
        """qudaHisqForce form: coeff[t] = (one-hop, three-hop) weights; with
        num_naik_terms > 0, len(multi_x) + num_naik_terms rows, the extra rows
        for the last num_naik_terms fields (second Naik set)."""
        self._activate()
        n = len(multi_x)
        if not 0 <= num_naik_terms <= n or len(coeff) != n + num_naik_terms:
            raise MilcHipError("hisq_force: coeff needs len(multi_x) + num_naik_terms rows")
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in multi_x]
        qptr = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        rows = [(ctypes.c_double * 2)(*c) for c in coeff]
        cptr = (ctypes.POINTER(ctypes.c_double) * len(rows))(*[ctypes.cast(r, ctypes.POINTER(ctypes.c_double))
                                                               for r in rows])
        l2 = (ctypes.c_double * 6)(*level2_coeff)
        l1 = (ctypes.c_double * 6)(*fat7_coeff)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (W, V, U)]
        _check(self.lib.milc_hip_hisq_force(2, n, num_naik_terms, eps, cptr, qptr, l2, l1, _p(arrs[0]), _p(arrs[1]),
                                            _p(arrs[2]), _p(mom)), "milc_hip_hisq_force")
        return mom

    def asqtad_force(self, act_path_coeff, one_link_oprod, naik_oprod, links, mom):
        """qudaAsqtadForce (generic_ks/fermion_force_asqtad_gpu.c:133-136): outer
        products (V, 4, 3, 3, 2) already scaled by 2 eps and the term weights;
        mom (V, 4, 10) updated in place."""
        self._activate()
        o1 = [np.ascontiguousarray(one_link_oprod[:, mu]) for mu in range(4)]
        o3 = [np.ascontiguousarray(naik_oprod[:, mu]) for mu in range(4)]
        d1 = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in o1])
        d3 = (ctypes.c_void_p * 4)(*[a.ctypes.data for a in o3])
        c = (ctypes.c_double * 6)(*act_path_coeff)
        U = np.ascontiguousarray(links, dtype=np.float64)
        _check(self.lib.milc_hip_asqtad_force(2, c, d1, d3, _p(U), _p(mom)), "milc_hip_asqtad_force")
        return mom

    # ---------------------------------------------------------------- device-resident API
    def dev_gauge_hisq(self, seed, scale=0.4):
        self._activate()
        nsvd = ctypes.c_int(0)
        _check(self.lib.milc_hip_dev_gauge_hisq(seed, scale, ctypes.byref(nsvd)), "milc_hip_dev_gauge_hisq")
        return nsvd.value

    def dev_source(self, seed, parity, slot):
        self._activate()
        _check(self.lib.milc_hip_dev_source(seed, parity, slot), "milc_hip_dev_source")

    def dev_invert(self, slot_b, slot_x, mass, parity, resid, mixed, max_iter=20000):
        self._activate()
        st = (ctypes.c_double * 6)()
        it = _check(self.lib.milc_hip_dev_invert(slot_b, slot_x, mass, parity, resid, mixed, max_iter, st),
                    "milc_hip_dev_invert")
        return it, {"resid": st[0], "seconds": st[1], "reliable_updates": int(st[2]), "converged": int(st[3]),
                    "iters": int(st[4]), "final_rsq": st[5]}

    def dev_block_tune(self, prec, nrhs):
        """(number of variants, chosen variant, its time in us, description)"""
        self._activate()
        v, us = ctypes.c_int(0), ctypes.c_double(0)
        buf = ctypes.create_string_buffer(256)
        n = _check(self.lib.milc_hip_dev_block_tune(prec, nrhs, ctypes.byref(v), ctypes.byref(us), buf, 256),
                   "milc_hip_dev_block_tune")
        return n, v.value, us.value, buf.value.decode()

    def dev_block_describe(self, prec, nrhs, variant):
        self._activate()
        buf = ctypes.create_string_buffer(256)
        _check(self.lib.milc_hip_dev_block_describe(prec, nrhs, variant, buf, 256), "milc_hip_dev_block_describe")
        return buf.value.decode()

    def dev_block_dslash_bench(self, prec, nrhs, variant, reps):
        self._activate()
        us = ctypes.c_double(0)
        _check(self.lib.milc_hip_dev_block_dslash_bench(prec, nrhs, variant, reps, ctypes.byref(us)),
               "milc_hip_dev_block_dslash_bench")
        return us.value

    def dev_block_dslash(self, prec, nrhs, variant, src_slots, out_parity, dst_slots):
        self._activate()
        n = len(src_slots)
        a = (ctypes.c_int * n)(*src_slots)
        b = (ctypes.c_int * n)(*dst_slots)
        _check(self.lib.milc_hip_dev_block_dslash(prec, nrhs, variant, n, a, out_parity, b), "milc_hip_dev_block_dslash")

    def dev_multishift_block(self, src_slots, offsets, resid, refine_resid, parity=EVEN, max_iters=20000,
                             out_slot0=-1, out_nsrc=None):
        self._activate()
        ns, n = len(src_slots), len(offsets)
        a = (ctypes.c_int * ns)(*src_slots)
        offs = (ctypes.c_double * n)(*offsets)
        st = (ctypes.c_double * (6 + 2 * n))()
        on = ns if out_nsrc is None else out_nsrc
        it = _check(self.lib.milc_hip_dev_multishift_block(ns, a, n, offs, resid, refine_resid, parity, max_iters,
                                                           out_slot0, on, st), "milc_hip_dev_multishift_block")
        return {"ms_iters": int(st[0]), "ms_seconds": st[1], "refine_seconds": st[2], "converged": int(st[3]),
                "variant": int(st[4]), "variant_usec": st[5], "true_rel": list(st[6:6 + n]),
                "refine_iters": [int(x) for x in st[6 + n:6 + 2 * n]], "iters": it}

    def dev_multishift(self, slot, offsets, resid, mixed, parity=EVEN, out_slot0=-1):
        """one source: (iterations, seconds, converged)"""
        self._activate()
        n = len(offsets)
        offs = (ctypes.c_double * n)(*offsets)
        st = (ctypes.c_double * 3)()
        it = _check(self.lib.milc_hip_dev_multishift(slot, parity, n, offs, resid, mixed, out_slot0, st),
                    "milc_hip_dev_multishift")
        return it, st[1], int(st[2])

    def dev_hmc_step(self, slot_phi, offsets, residues, resid, mixed=1, compare_fp64=True, thin_force=False):
        self._activate()
        n = len(offsets)
        offs = (ctypes.c_double * n)(*offsets)
        res = (ctypes.c_double * n)(*residues)
        st = (ctypes.c_double * 13)()
        _check(self.lib.milc_hip_dev_hmc_step(slot_phi, n, offs, res, resid, mixed, int(bool(compare_fp64)),
                                              int(bool(thin_force)), st), "milc_hip_dev_hmc_step")
        return {"force_seconds": st[0], "force_iters": int(st[1]), "S": st[2], "action": st[3],
                "action_seconds": st[4], "action_iters": int(st[5]), "rel_err_O1": st[6], "rel_err_O3": st[7],
                "fp64_force_seconds": st[8], "thin_force_seconds": st[9], "rel_err_mom": st[10],
                "action_fp64": st[11], "rel_err_action": st[12]}

    def dev_hisq_force(self, slot0, residues, eps, mom=None):
        """HISQ thin-link force from the even-site solutions in slots slot0..;
        adds into mom ((V, 4, 10) FP64, MILC order) when given"""
        self._activate()
        n = len(residues)
        res = (ctypes.c_double * n)(*residues)
        st = (ctypes.c_double * 4)()
        mp = None
        if mom is not None:
            assert mom.dtype == np.float64 and mom.flags.c_contiguous and mom.shape == (self.V, 4, 10)
            mp = _p(mom)
        _check(self.lib.milc_hip_dev_hisq_force(slot0, n, res, eps, mp, st), "milc_hip_dev_hisq_force")
        return {"oprod_seconds": st[0], "thin_force_seconds": st[1], "mom_norm2": st[2], "aux_seconds": st[3]}

    def dev_read(self, slot, parity):
        self._activate()
        out = np.zeros((self.V, 3, 2))
        _check(self.lib.milc_hip_dev_read(slot, parity, _p(out)), "milc_hip_dev_read")
        return out

    def dev_write(self, slot, parity, v):
        self._activate()
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.lib.milc_hip_dev_write(slot, parity, _p(v)), "milc_hip_dev_write")

    def dev_load_thin_links(self, thin, phases_in=False):
        """thin links (V, 4, 3, 3, 2) in MILC layout -> device, HISQ links built
        there; returns the number of links that took the SVD fallback"""
        self._activate()
        U = np.ascontiguousarray(thin, dtype=np.float64)
        nsvd = ctypes.c_int(0)
        _check(self.lib.milc_hip_dev_load_thin_links(2, _p(U), int(bool(phases_in)), ctypes.byref(nsvd)),
               "milc_hip_dev_load_thin_links")
        return nsvd.value

    def dev_read_thin_links(self, with_phases=False):
        self._activate()
        U = np.zeros((self.V, 4, 3, 3, 2))
        _check(self.lib.milc_hip_dev_read_thin_links(_p(U), int(bool(with_phases))), "milc_hip_dev_read_thin_links")
        return U

    def dev_read_links(self):
        self._activate()
        fat = np.zeros((self.V, 4, 3, 3, 2))
        lng = np.zeros((self.V, 4, 3, 3, 2))
        _check(self.lib.milc_hip_dev_read_links(_p(fat), _p(lng)), "milc_hip_dev_read_links")
        return fat, lng

    def dev_dslash(self, prec, slot_src, parity_out, slot_dst):
        """slot_dst = D slot_src on parity_out, stencil in precision prec (0/1/2)"""
        self._activate()
        _check(self.lib.milc_hip_dev_dslash(prec, slot_src, parity_out, slot_dst), "milc_hip_dev_dslash")

    def dev_dslash_bench(self, prec, parity, reps):
        self._activate()
        us = ctypes.c_double(0)
        _check(self.lib.milc_hip_dev_dslash_bench(prec, parity, reps, ctypes.byref(us)), "milc_hip_dev_dslash_bench")
        return us.value

    def dev_cg_iteration_bench(self, slot, mass, parity=EVEN, reps=50):
        """16-bit mixed CG iteration kernels timed in their own sequence (us):
        dslash, dslash + M^dag M dots, scalars, update, whole iteration"""
        self._activate()
        us = (ctypes.c_double * 5)()
        _check(self.lib.milc_hip_dev_cg_iteration_bench(slot, mass, parity, reps, us),
               "milc_hip_dev_cg_iteration_bench")
        return {"dslash": us[0], "dslash_mdagm": us[1], "cg_scalars": us[2], "cg_update": us[3],
                "iteration": us[4]}

    def dev_copy_bench(self, nbytes, reps):
        self._activate()
        us = ctypes.c_double(0)
        _check(self.lib.milc_hip_dev_copy_bench(nbytes, reps, ctypes.byref(us)), "milc_hip_dev_copy_bench")
        return us.value

    def dev_diff(self, slot_a, slot_b):
        self._activate()
        r = ctypes.c_double(0)
        _check(self.lib.milc_hip_dev_diff(slot_a, slot_b, ctypes.byref(r)), "milc_hip_dev_diff")
        return r.value

    def dev_load_links(self, fat, lng):
        """Make host fat/long links (MILC layout) the current device links."""
        self._activate()
        fat = np.ascontiguousarray(fat, dtype=np.float64)
        lng = np.ascontiguousarray(lng, dtype=np.float64)
        _check(self.lib.milc_hip_dev_load_links(_p(fat), _p(lng)), "milc_hip_dev_load_links")
        self._fresh = True

    def dev_rhmc_action(self, slot_phi, roots, residues, resid, mixed=0):
        """fermion_action: sum_even |a0 phi + sum_k a_k x_k|^2 (roots[0] unused)."""
        self._activate()
        order = len(roots) - 1
        r = (ctypes.c_double * (order + 1))(*roots)
        a = (ctypes.c_double * (order + 1))(*residues)
        out = ctypes.c_double(0)
        it = _check(self.lib.milc_hip_dev_rhmc_action(slot_phi, order, r, a, resid, mixed, ctypes.byref(out)),
                    "milc_hip_dev_rhmc_action")
        return out.value, it

    def dev_force_level0(self, slot_phi, offsets, residues, resid, mixed=0, want_oprod=True):
        """(S, O1, O3, iterations) of S = sum_t r_t phi^dag (M^dag M + sigma_t)^-1 phi."""
        self._activate()
        n = len(offsets)
        o = (ctypes.c_double * n)(*offsets)
        r = (ctypes.c_double * n)(*residues)
        S = ctypes.c_double(0)
        o1 = np.zeros((self.V, 4, 3, 3, 2)) if want_oprod else None
        o3 = np.zeros((self.V, 4, 3, 3, 2)) if want_oprod else None
        it = _check(self.lib.milc_hip_dev_force_level0(slot_phi, n, o, r, resid, mixed, ctypes.byref(S),
                                                       _p(o1) if want_oprod else None,
                                                       _p(o3) if want_oprod else None),
                    "milc_hip_dev_force_level0")
        return S.value, o1, o3, it

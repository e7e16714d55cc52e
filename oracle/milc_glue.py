"""MILC's own solver and force callers through the glue of integration/.

TEST INFRASTRUCTURE ONLY.  Makefile.glue builds MILC 7.8.1's CPU code with
glue_driver.c in four variants (MILC_PRECISION 1 / 2, MILC's CPU build / the
-DUSE_CG_GPU build whose offload goes through integration/*.c into
libmilc_hip.so).  Run as a script, this file loads the CPU and the HIP variant
of one precision into one process (each keeps its own MILC globals), feeds
both the same thin links, lets MILC build its HISQ links in each, and calls
ks_congrad_field, ks_multicg_field and eo_fermion_force_multi through both;
it writes the comparisons as JSON.  It runs in a process of its own because
MILC's lattice is a set of process globals and the HIP variant holds the
library's device context.

    python oracle/milc_glue.py --precision 2 --out result.json
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EVEN, ODD, EVENANDODD = 2, 1, 3


def lib_path(variant):
    return os.path.join(HERE, "_ref", "glue", variant, "libmilc_glue.so")


def available():
    return all(os.path.exists(lib_path(v)) for v in ("cpu_p1", "cpu_p2", "hip_p1", "hip_p2"))


class Build:
    """One variant of the harness: MILC's callers in one build."""

    def __init__(self, variant):
        L = ctypes.CDLL(lib_path(variant))
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.drv_init.argtypes = [I, I, I, I]
        L.drv_sites.restype = ctypes.c_size_t
        L.drv_set_thin_links.argtypes = [P]
        L.drv_hisq_links.argtypes = [P, P]
        L.drv_congrad_field.argtypes = [P, P, I, D, I, I, D, D, I, P]
        L.drv_multicg_field.argtypes = [P, ctypes.POINTER(P), P, I, I, I, D, I, P]
        L.drv_force.argtypes = [ctypes.POINTER(P), P, I, D, P]
        self.lib = L
        self.real = np.float32 if L.drv_real_bytes() == 4 else np.float64
        self.prec = 1 if self.real == np.float32 else 2

    def init(self, dims):
        if self.lib.drv_init(*dims) != 0:
            raise RuntimeError("drv_init failed")
        self.V = int(self.lib.drv_sites())

    def set_thin_links(self, U):
        self.lib.drv_set_thin_links(np.ascontiguousarray(U, dtype=np.float64).ctypes.data)

    def hisq_links(self):
        fat = np.zeros((self.V, 4, 3, 3, 2))
        lng = np.zeros((self.V, 4, 3, 3, 2))
        self.lib.drv_hisq_links(fat.ctypes.data, lng.ctypes.data)
        return fat, lng

    def congrad_field(self, src, dest, parity, mass, niter, nrestart, resid, relresid=0.0):
        src = np.ascontiguousarray(src, dtype=self.real)
        dest = np.ascontiguousarray(dest, dtype=self.real)
        out = np.zeros(5)
        it = self.lib.drv_congrad_field(src.ctypes.data, dest.ctypes.data, parity, mass, niter, nrestart, resid,
                                        relresid, self.prec, out.ctypes.data)
        return dest, it, {"final_rsq": out[0], "final_relrsq": out[1], "converged": int(out[2]),
                          "final_iters": int(out[3]), "final_restart": int(out[4])}

    def multicg_field(self, src, offsets, parity, niter, resid):
        src = np.ascontiguousarray(src, dtype=self.real)
        n = len(offsets)
        psim = [np.zeros_like(src) for _ in range(n)]
        ptrs = (ctypes.c_void_p * n)(*[p.ctypes.data for p in psim])
        offs = np.ascontiguousarray(offsets, dtype=np.float64)
        out = np.zeros(n)
        it = self.lib.drv_multicg_field(src.ctypes.data, ptrs, offs.ctypes.data, n, parity, niter, resid, self.prec,
                                        out.ctypes.data)
        return psim, it, out

    def force(self, multi_x, residues, eps):
        xs = [np.ascontiguousarray(x, dtype=self.real) for x in multi_x]
        n = len(xs)
        ptrs = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        res = np.ascontiguousarray(residues, dtype=np.float64)
        mom = np.zeros((self.V, 4, 10))
        if self.lib.drv_force(ptrs, res.ctypes.data, n, eps, mom.ctypes.data) != 0:
            raise RuntimeError("drv_force failed")
        return mom


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(precision, dims=(8, 8, 8, 8), log=print):
    """Both builds of one precision on the same inputs; returns a dict of
    comparisons."""
    sys.path.insert(0, HERE)
    import ks_oracle as ko

    cpu, hip = Build(f"cpu_p{precision}"), Build(f"hip_p{precision}")
    for b in (cpu, hip):
        b.init(dims)
    L = ko.Lattice(dims)
    U = L.gauge_random(1000)
    for b in (cpu, hip):
        b.set_thin_links(U)
    fc, lc = cpu.hisq_links()
    fh, lh = hip.hisq_links()
    out = {"precision": precision, "dims": list(dims),
           "links_cpu_vs_hip_build": max(rel(fh, fc), rel(lh, lc))}
    s_even, s_odd = L.parity_slice(ko.EVEN), L.parity_slice(ko.ODD)
    resid = 1e-10 if precision == 2 else 2e-6

    def solve_pair(tag, src, parity, mass, resid, relresid=0.0, niter=500, nrestart=5):
        t0 = time.time()
        xc, ic, qc = cpu.congrad_field(src, np.zeros_like(src), parity, mass, niter, nrestart, resid, relresid)
        t1 = time.time()
        xh, ih, qh = hip.congrad_field(src, np.zeros_like(src), parity, mass, niter, nrestart, resid, relresid)
        t2 = time.time()
        sl = {EVEN: s_even, ODD: s_odd, EVENANDODD: slice(None)}[parity]
        out[tag] = {"iters_cpu": ic, "iters_hip": ih, "qic_cpu": qc, "qic_hip": qh,
                    "rel_solution": rel(xh[sl], xc[sl]),
                    "untouched_other_parity": bool(parity == EVENANDODD or
                                                   not np.any(xh[s_odd if parity == EVEN else s_even])),
                    "sec_cpu": t1 - t0, "sec_hip": t2 - t1}
        log(f"{tag}: iters cpu {ic} hip {ih}; rel {out[tag]['rel_solution']:.3e}; "
            f"final_rsq cpu {qc['final_rsq']:.3e} hip {qh['final_rsq']:.3e}; "
            f"final_relrsq cpu {qc['final_relrsq']:.3e} hip {qh['final_relrsq']:.3e}")

    b_even = L.gaussian_vector(7, ko.EVEN)
    b_all = L.gaussian_vector(9, ko.EVENANDODD)
    solve_pair("cg_even", b_even, EVEN, 0.05, resid)
    solve_pair("cg_evenandodd", b_all, EVENANDODD, 0.1, resid)
    # Fermilab relative residual: alone (resid = 0) and beside the residual
    relres = 1e-4 if precision == 2 else 3e-3
    solve_pair("cg_relresid_only", b_even, EVEN, 0.05, 0.0, relres)
    solve_pair("cg_relresid_and_resid", b_even, EVEN, 0.05, resid * 10, relres)

    offs = [4 * 0.05 ** 2 + x for x in (0.0, 0.02, 0.2, 1.0)]
    mresid = 1e-10 if precision == 2 else 1e-5
    pc, itc, rc_ = cpu.multicg_field(b_even, offs, EVEN, 2000, mresid)
    ph, ith, rh = hip.multicg_field(b_even, offs, EVEN, 2000, mresid)
    out["multicg"] = {"iters_cpu": itc, "iters_hip": ith, "final_rsq_cpu": list(rc_), "final_rsq_hip": list(rh),
                      "rel_solution": [rel(ph[j][s_even], pc[j][s_even]) for j in range(len(offs))]}
    log(f"multicg: iters cpu {itc} hip {ith}; rel {out['multicg']['rel_solution']}")

    xs = [L.gaussian_vector(100 + t, ko.EVENANDODD) for t in range(3)]
    res = [0.3, -1.2, 0.7]
    mc = cpu.force(xs, res, 0.02)
    mh = hip.force(xs, res, 0.02)
    out["force"] = {"rel_mom": rel(mh, mc), "norm": float(np.linalg.norm(mc))}
    log(f"force: rel {out['force']['rel_mom']:.3e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", type=int, default=2)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    res = run(a.precision, log=lambda s: print(s, flush=True))
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()

"""MILC's CPU side of the 32^3x64 comparison (run as a subprocess by
tests/test_gpu_c2_vs_milc.py; not collected by pytest).

MILC's lattice lives in process globals and the test process already holds an
8^4 MILC lattice, so the full-size MILC run gets a process of its own.  Inputs
(numpy files in --dir, written by the test): U (thin links, no phases), fat /
lng (the device's HISQ links), b (source, even sites), x64 / xmix (the device's
FP64 and mixed solutions).  MILC builds its own HISQ links from U
(fermion_links_hisq_load_milc.c), solves with ks_congrad_parity_cpu
(d_congrad5_fn_milc.c:60-407) and recomputes every residual with its own
dslash_fn_field (dslash_fn_dblstore.c).  Writes result.json and x_milc.npy.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import milc_ref as mr  # noqa: E402


def log(s):
    print(s, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True)
    ap.add_argument("--dims", default="32,32,32,64")
    ap.add_argument("--mass", type=float, default=0.01)
    ap.add_argument("--resid", type=float, default=1e-14)
    ap.add_argument("--niter", type=int, default=1000)
    ap.add_argument("--nrestart", type=int, default=5)
    a = ap.parse_args()
    dims = tuple(int(x) for x in a.dims.split(","))
    d = a.dir
    ld = lambda n: np.load(os.path.join(d, n + ".npy"), mmap_mode="r")  # noqa: E731
    out = {}
    t0 = time.time()
    V = mr.init(dims)
    Vh = V // 2
    mr.set_thin_links(np.asarray(ld("U")))
    fat, lng = mr.hisq_links(V)
    out["milc_links_seconds"] = time.time() - t0
    gfat, glng = ld("fat"), ld("lng")

    def rel(x, y):
        return float(np.linalg.norm(np.asarray(x) - np.asarray(y)) / np.linalg.norm(np.asarray(y)))

    out["rel_fat_gpu_vs_milc"] = rel(gfat, fat)
    out["rel_lng_gpu_vs_milc"] = rel(glng, lng)
    log(f"links: MILC built them in {out['milc_links_seconds']:.1f} s; device vs MILC fat "
        f"{out['rel_fat_gpu_vs_milc']:.2e} long {out['rel_lng_gpu_vs_milc']:.2e}")
    del fat, lng
    b = np.ascontiguousarray(ld("b"))
    msq4 = 4.0 * a.mass * a.mass

    def true_resid(x):
        # |b - M^dag M x| / |b| on even sites, M^dag M = 4m^2 - D_eo D_oe
        t = mr.dslash(x, mr.ODD)
        u = mr.dslash(t, mr.EVEN)
        r = b[:Vh] - (msq4 * x[:Vh] - u[:Vh])
        return float(np.linalg.norm(r) / np.linalg.norm(b[:Vh]))

    t0 = time.time()
    x, it, q = mr.congrad(b, a.mass, a.resid, mr.EVEN, a.niter, a.nrestart)
    out["milc_seconds"] = time.time() - t0
    out["milc_iters"] = it
    out["milc_final_rsq"] = q["final_rsq"]
    out["milc_restarts"] = q["restarts"]
    out["milc_converged"] = q["converged"]
    out["true_resid_milc"] = true_resid(x)
    log(f"MILC CG: {it} iterations in {out['milc_seconds']:.1f} s, true residual {out['true_resid_milc']:.3e}")
    for name in ("x64", "xmix"):
        y = np.ascontiguousarray(ld(name))
        out["true_resid_" + name] = true_resid(y)
        out["rel_" + name + "_vs_milc"] = rel(y[:Vh], x[:Vh])
        log(f"{name}: true residual (MILC's Dslash) {out['true_resid_' + name]:.3e}, "
            f"relative difference from MILC's solution {out['rel_' + name + '_vs_milc']:.3e}")
    np.save(os.path.join(d, "x_milc.npy"), x)
    with open(os.path.join(d, "result.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

"""CPU-side checks of the product's host pieces (no GPU needed).

* the C-ABI library loads and exports every symbol include/milc_hip.h declares;
* the device stencil indexing (mh_common.h site_coords/neighbour), compiled
  for the host, reproduces the oracle's MILC neighbour tables on every site.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import ks_oracle as ko

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    """every function declared in include/*.h (milc_hip_* and milc_io_*)"""
    syms = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        syms |= set(re.findall(r"\b(milc_(?:hip|io)_[a-z_0-9]+)\s*\(", txt))
    return sorted(syms)


def test_library_exports_every_header_symbol():
    import milc_qcd_amd as mq
    lib_path = mq.LIB_PATH
    assert os.path.exists(lib_path), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    syms = header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert set(mq.EXPORTS) <= set(syms)


def test_library_loads_without_gpu():
    import milc_qcd_amd as mq
    lib = mq.load_library()
    assert b"gfx950" in lib.milc_hip_version()


def test_fastdiv_magic_numbers():
    # the Granlund-Montgomery constants used by mh_common.h FastDiv
    rng = np.random.default_rng(5)
    for d in list(range(1, 130)) + list(rng.integers(1, 1 << 20, 50)):
        d = int(d)
        s = 0
        while (1 << s) < d:
            s += 1
        m = (((1 << 32) * ((1 << s) - d)) // d + 1) & 0xFFFFFFFF
        for n in [0, 1, d - 1, d, d + 1, (1 << 31) - 1] + list(rng.integers(0, 1 << 31, 50)):
            n = int(n)
            assert (((n * m) >> 32) + n) >> s == n // d


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 4, 8), (8, 4, 6, 4), (6, 4, 4, 6)])
def test_device_indexing_matches_milc_neighbours(dims, tmp_path):
    exe = tmp_path / "geom_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-o", str(exe),
                    os.path.join(ROOT, "tests", "geom_check.cpp")], check=True)
    out = subprocess.run([str(exe)] + [str(d) for d in dims], capture_output=True, text=True, check=True).stdout
    tab = np.array([[int(v) for v in l.split()] for l in out.splitlines()])
    L = ko.Lattice(dims)
    Vh = L.Vh
    nbr = np.ctypeslib.as_array(L._g.nbr, shape=(16, L.V))
    for row in tab:
        pb, cb = row[0], row[1]
        i = pb * Vh + cb
        assert tuple(row[2:6]) == tuple(L.coords[i])
        nb = row[6:].reshape(4, 4)  # mu x (+1, +3, -1, -3)
        for mu in range(4):
            expect = [nbr[mu][i], nbr[8 + mu][i], nbr[7 - mu][i], nbr[15 - mu][i]]
            got = [(1 - pb) * Vh + int(v) for v in nb[mu]]
            assert got == list(expect), (dims, pb, cb, mu)


@pytest.mark.parametrize("dims,block", [((24, 24, 24, 48), (8, 8, 8, 4)), ((8, 6, 4, 10), (4, 2, 2, 2)),
                                        ((16, 16, 16, 32), (8, 4, 4, 4)), ((8, 8, 8, 8), (4, 4, 4, 4))])
def test_blocked_site_order_is_a_bijection(tmp_path, dims, block):
    """the batched stencil's 4-D blocked site order visits every checkerboard
    site of each parity exactly once, with the coordinates site_coords gives"""
    exe = tmp_path / "border_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-o", str(exe),
                    os.path.join(ROOT, "tests", "border_check.cpp")], check=True)
    out = subprocess.run([str(exe)] + [str(d) for d in dims] + [str(b) for b in block], capture_output=True,
                         text=True, check=True).stdout
    assert out.strip() == "OK", out

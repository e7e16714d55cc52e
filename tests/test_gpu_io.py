"""Lattice files feeding the device path: a configuration generated on the
device is read back without its KS phases, written with the library's
writers (v5, SciDAC, ILDG), read back, loaded onto the device and smeared
there; the HISQ links match the oracle's smearing of the same thin links and,
for a 64-bit file, the links built from the seed bit for bit.

Workload: 6x4x8x10 lattice, gauge seed 1040 (scale 0.4); tolerance 1e-12 for
the smearing against the oracle.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq
from milc_qcd_amd import lattice_io as lio

pytestmark = pytest.mark.gpu

DIMS = (6, 4, 8, 10)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_thin_links_round_trip_through_files(tmp_path):
    L = ko.Lattice(DIMS)
    U_bare = L.gauge_random(1040)
    fat_ref, lng_ref, _ = L.hisq_links(L.rephase(U_bare))
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1040, 0.4)
        fat0, lng0 = G.dev_read_links()
        U = G.dev_read_thin_links(with_phases=False)
        assert rel(U, U_bare) < 1e-14
        assert rel(G.dev_read_thin_links(with_phases=True), L.rephase(U_bare)) < 1e-14
        for flag, prec in ((lio.SAVE_SERIAL_SCIDAC, 2), (lio.SAVE_SERIAL_ILDG, 2), (lio.SAVE_SERIAL, 1),
                           (lio.SAVE_SERIAL_SCIDAC, 1)):
            f = str(tmp_path / f"cfg_{flag}_{prec}")
            lio.save_lattice(flag, f, U, list(DIMS), precision=prec)
            U2, info = lio.reload_lattice(f, list(DIMS))
            assert info["checksum_ok"]
            G.dev_load_thin_links(U2)
            fat, lng = G.dev_read_links()
            if prec == 2:
                assert np.array_equal(fat, fat0) and np.array_equal(lng, lng0)
            # the device smearing of the file's links against the oracle's
            fat_o, lng_o, _ = L.hisq_links(L.rephase(U2))
            assert rel(fat, fat_o) < 1e-12 and rel(lng, lng_o) < 1e-12
            # the FP32 file round trip keeps the links to FP32 accuracy
            assert rel(fat, fat_ref) < (1e-14 if prec == 2 else 1e-6)


def test_solver_on_links_from_file(tmp_path):
    """a CG on links loaded from a SciDAC file reaches its FP64 target and
    matches the oracle's solve on the same links"""
    L = ko.Lattice(DIMS)
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1041, 0.4)
        U = G.dev_read_thin_links()
        f = str(tmp_path / "cfg")
        lio.save_lattice(lio.SAVE_SERIAL_SCIDAC, f, U, list(DIMS), precision=1)
        U2, _ = lio.reload_lattice(f)
        G.dev_load_thin_links(U2)
        fat, lng, _ = L.hisq_links(L.rephase(U2))
        src = L.gaussian_vector(1042, ko.EVEN)
        G.dev_write(0, ko.EVEN, src)
        it, st = G.dev_invert(0, 1, 0.05, mq.EVEN, 1e-10, 2)
        x = G.dev_read(1, ko.EVEN)
    ref, _, _ = L.congrad(fat, lng, src, 0.05, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=5)
    sl = L.parity_slice(ko.EVEN)
    assert st["converged"] == 1 and st["resid"] < 1e-10
    assert rel(x[sl], ref[sl]) < 1e-8

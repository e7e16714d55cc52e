"""Lattice-file readers and writers (include/milc_io.h, host code) against the
reference's own files and numbers:

* the oracle (oracle/ks_io.py) reproduces the checksums stored in the files
  and those MILC printed on reading them, and MILC's printed plaquette and
  link trace (which fix site and direction order);
* the library reads the same configuration from v5 (both byte orders),
  SciDAC and ILDG files bit-identically to the oracle, in node_index order,
  and reports the stored/computed checksums the same way;
* writer -> reader round trips in every format and precision, the writers'
  checksums agree with the oracle's, and error handling (truncated file,
  wrong dims, foreign magic, corrupted payload) returns the documented codes.

Fixtures: tests/golden/io/ (MILC's binary_samples files, unchanged).
"""
import os
import shutil

import numpy as np
import pytest

import ks_io as oio
import milc_qcd_amd as mq
from milc_qcd_amd import lattice_io as lio

HERE = os.path.dirname(os.path.abspath(__file__))
IO = os.path.join(HERE, "golden", "io")

# MILC's printed values on reading these files (*.sample-out of the reference):
# hvy_qpot/test/su3_hqp.1 (l4444), wilson_flow/test/wilson_flow.1 (l4448),
# hvy_qpot/test/su3_hybrid_hqp_hyp.level1.2 (l6666),
# schroed_cl_inv/test/su3_schr_cl_bi.1 (5p7: stored values of a file whose
# stored checksum does not match its data)
PINS = {
    "lat.sample.l4444": {"sums": (0x2352C05, 0xD137321D), "plaq": (1.794675, 1.774426), "linktr": 6.467587e-01},
    "lat.sample.l4448": {"sums": (0x13F3B413, 0x161F7DDE), "plaq": (1.723748, 1.690586), "linktr": 6.921659e-02},
    "lat.sample.l6666": {"sums": (0xC1D08F5, 0x68164BEF), "plaq": (1.9827179876982368, 1.9811715330156219),
                         "linktr": 9.0159201231658637e-01},
}


def path(name):
    return os.path.join(IO, name)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", sorted(PINS))
def test_oracle_v5_pins(name):
    r = oio.read_v5(path(name))
    p = PINS[name]
    assert (r["sum29"], r["sum31"]) == p["sums"]
    assert (r["calc29"], r["calc31"]) == p["sums"]
    L = r["links"].astype(np.float64)
    ss, st = oio.plaquette(L, r["dims"])
    assert abs(ss - p["plaq"][0]) < 2e-6 * abs(p["plaq"][0]) and abs(st - p["plaq"][1]) < 2e-6 * abs(p["plaq"][1])
    assert abs(oio.link_trace(L) - p["linktr"]) < 2e-6 * abs(p["linktr"])


def test_oracle_lime_checksums_and_cross_format():
    for base in ("lat.sample.l4444", "lat.sample.l4448"):
        v5 = oio.read_v5(path(base))["links"].astype(np.float64)
        for ext in (".scidac", ".ildg", ".foreign"):
            if not os.path.exists(path(base + ext)):
                continue
            f = oio.read_lime_field(path(base + ext))
            if f["suma"] is not None:
                assert (f["calca"], f["calcb"]) == (f["suma"], f["sumb"])
            assert np.array_equal(f["data"], v5)
    f = oio.read_lime_field(path("ksvector.src.sample.l8"))
    assert (f["calca"], f["calcb"]) == (f["suma"], f["sumb"]) == (0x18C175B6, 0xADC0F9EE)


def test_oracle_ksvector_matches_golden_fixture():
    f = oio.read_lime_field(path("ksvector.src.sample.l8"))
    gold = np.load(os.path.join(HERE, "golden", "ksvector_src_l8.npy"))
    assert np.array_equal(f["data"][:, 0], gold.astype(np.float64))


def to_node_index(lex_field, dims):
    out = np.empty_like(lex_field)
    out[oio.node_index_of_lex(dims)] = lex_field
    return out


@pytest.mark.parametrize("name", ["lat.sample.l4444", "lat.sample.l4444.scidac", "lat.sample.l4444.ildg",
                                  "lat.sample.l4448", "lat.sample.l4448.scidac", "lat.sample.l4448.foreign",
                                  "lat.sample.l6666"])
def test_library_reads_gauge_files(name):
    base = name.split(".scidac")[0].split(".ildg")[0].split(".foreign")[0]
    ref = oio.read_v5(path(base))
    want = to_node_index(ref["links"].astype(np.float64), ref["dims"])
    links, info = lio.reload_lattice(path(name))
    assert info["dims"] == ref["dims"]
    assert np.array_equal(links, want)
    if name in PINS:
        assert info["format"] == "v5" and info["checksum_ok"]
        assert info["stored"] == info["computed"] == PINS[name]["sums"]
        assert info["byte_reversed"] == ref["byte_reversed"]
    elif name.endswith(".foreign"):
        assert info["format"] == "ildg" and not info["has_checksum"]
    else:
        f = oio.read_lime_field(path(name))
        assert info["checksum_ok"] and info["stored"] == (f["suma"], f["sumb"]) == info["computed"]


def test_library_reports_v5_checksum_violation():
    """lat.sample.5p7: stored checksum does not match the data; MILC reads it
    with a warning and so does the library (return 0, checksum_ok False)"""
    r = oio.read_v5(path("lat.sample.5p7"))
    links, info = lio.reload_lattice(path("lat.sample.5p7"))
    assert not info["checksum_ok"]
    assert info["stored"] == (0x2E218AB1, 0x2D188E6B) == (r["sum29"], r["sum31"])
    assert info["computed"] == (r["calc29"], r["calc31"])
    assert np.array_equal(links, to_node_index(r["links"].astype(np.float64), r["dims"]))


def test_library_reads_ksvector():
    dims = [8, 8, 8, 8]
    v, info = lio.restore_ks_vector_scidac_to_field(path("ksvector.src.sample.l8"), dims)
    gold = np.load(os.path.join(HERE, "golden", "ksvector_src_l8.npy")).astype(np.float64)
    assert np.array_equal(v, to_node_index(gold, dims))
    assert info["checksum_ok"] and info["stored"] == (0x18C175B6, 0xADC0F9EE)


@pytest.fixture
def field():
    dims = [4, 6, 4, 8]
    rng = np.random.default_rng(11)
    V = int(np.prod(dims))
    return dims, rng.standard_normal((V, 4, 3, 3, 2)), rng.standard_normal((V, 3, 2))


@pytest.mark.parametrize("flag,prec", [(lio.SAVE_SERIAL, 1), (lio.SAVE_SERIAL_SCIDAC, 1),
                                       (lio.SAVE_SERIAL_SCIDAC, 2), (lio.SAVE_SERIAL_ILDG, 1),
                                       (lio.SAVE_SERIAL_ILDG, 2)])
def test_gauge_round_trip(tmp_path, field, flag, prec):
    dims, U, _ = field
    f = str(tmp_path / "cfg")
    winfo = lio.save_lattice(flag, f, U, dims, precision=prec)
    back, info = lio.reload_lattice(f, dims)
    expect = U.astype(np.float32).astype(np.float64) if prec == 1 else U
    assert np.array_equal(back, expect)
    assert info["checksum_ok"] and info["stored"] == winfo["stored"]
    # the oracle reads the library's file and agrees on data and checksums
    lex = U[oio.node_index_of_lex(dims)]
    lexexp = lex.astype(np.float32).astype(np.float64) if prec == 1 else lex
    if flag == lio.SAVE_SERIAL:
        r = oio.read_v5(f)
        assert (r["sum29"], r["sum31"]) == (r["calc29"], r["calc31"]) == winfo["stored"]
        assert np.array_equal(r["links"].astype(np.float64), lexexp)
    else:
        r = oio.read_lime_field(f)
        assert (r["suma"], r["sumb"]) == (r["calca"], r["calcb"]) == winfo["stored"]
        assert np.array_equal(r["data"], lexexp)


@pytest.mark.parametrize("prec", [1, 2])
def test_vector_and_matrix_round_trip(tmp_path, field, prec):
    dims, U, v = field
    f = str(tmp_path / "vec")
    lio.save_ks_vector_scidac_from_field(f, v, dims, precision=prec)
    back, info = lio.restore_ks_vector_scidac_to_field(f, dims)
    assert np.array_equal(back, v.astype(np.float32).astype(np.float64) if prec == 1 else v)
    assert info["datatype"] == "ColorVector" and info["checksum_ok"]
    f2 = str(tmp_path / "mom")
    lio.save_color_matrix_scidac_from_field(f2, U, dims, 4, precision=prec)
    m, info = lio.restore_color_matrix_scidac_to_field(f2, dims, 4)
    assert np.array_equal(m, U.astype(np.float32).astype(np.float64) if prec == 1 else U)
    # a vector file is not a gauge file and vice versa
    with pytest.raises(mq.MilcHipError):
        lio.reload_lattice(f, dims)
    with pytest.raises(mq.MilcHipError):
        lio.restore_ks_vector_scidac_to_field(f2, dims)


def status(fn):
    try:
        fn()
    except mq.MilcHipError as e:
        return int(str(e).split("status")[1])
    return 0


def test_error_codes(tmp_path):
    dims = [4, 4, 4, 4]
    assert status(lambda: lio.reload_lattice(str(tmp_path / "missing"), dims)) == -1
    # truncated v5 and LIME files
    for name in ("lat.sample.l4444", "lat.sample.l4444.scidac"):
        data = open(path(name), "rb").read()
        t = tmp_path / ("trunc_" + name)
        t.write_bytes(data[:len(data) // 2])
        assert status(lambda: lio.reload_lattice(str(t), dims)) in (-1, -2)
    # wrong dimensions
    assert status(lambda: lio.reload_lattice(path("lat.sample.l4444"), [4, 4, 4, 8])) == -4
    # not a lattice file
    junk = tmp_path / "junk"
    junk.write_bytes(b"\x00\x01\x02\x03" * 100)
    assert status(lambda: lio.reload_lattice(str(junk), dims)) == -2
    # a flipped payload byte in a SciDAC file: the checksum fails the read (-3)
    data = bytearray(open(path("lat.sample.l4444.scidac"), "rb").read())
    k = data.find(b"scidac-binary-data") + 128 + 1000
    data[k] ^= 0x40
    bad = tmp_path / "bad.scidac"
    bad.write_bytes(bytes(data))
    assert status(lambda: lio.reload_lattice(str(bad), dims)) == -3
    info = lio.read_info(str(bad))
    assert info["has_checksum"] and not info["checksum_ok"]

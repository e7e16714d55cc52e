"""Extract MILC's own golden test data into small .npy fixtures.

Reads the SciDAC/LIME files that MILC's check_ks_invert / check_fermion_force
tests use (binary_samples/*.l8 in the reference tree) and writes their binary
payloads, converted from big-endian float32 to native float32 in lexicographic
site order (x fastest), as .npy files next to this script.  The files are data
(inputs and expected outputs of MILC's own tests); no reference code is kept.

LIME record layout: 144-byte header (u32 magic 0x456789ab, u16 version,
u16 flags, u64 payload length, 128-byte type string), payload padded to 8 bytes.

The lattice files used by the file-format tests are copied unchanged into
tests/golden/io/.

Usage: python tests/golden/make_golden_fixtures.py [/root/reference]
"""
import os
import re
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_lime(path):
    data = open(path, "rb").read()
    off, recs = 0, []
    while off < len(data):
        magic, _ver, _flags, length = struct.unpack(">IHHQ", data[off:off + 16])
        if magic != 0x456789AB:
            raise ValueError(f"{path}: bad LIME magic at {off}")
        typ = data[off + 16:off + 144].split(b"\0")[0].decode()
        recs.append((typ, data[off + 144:off + 144 + length]))
        off += 144 + ((length + 7) // 8) * 8
    return recs


def scidac_field(path):
    recs = read_lime(path)
    meta = dict(recs)
    dims = [int(x) for x in re.search(rb"<dims>([^<]*)</dims>", meta["scidac-private-file-xml"]).group(1).split()]
    rx = meta["scidac-private-record-xml"]
    dtype = re.search(rb"<datatype>([^<]*)</datatype>", rx).group(1).decode()
    prec = re.search(rb"<precision>([^<]*)</precision>", rx).group(1).decode()
    count = int(re.search(rb"<datacount>([^<]*)</datacount>", rx).group(1))
    if prec != "F":
        raise ValueError("only single-precision fixtures are expected")
    raw = np.frombuffer(meta["scidac-binary-data"], dtype=">f4").astype(np.float32)
    V = int(np.prod(dims))
    if dtype == "QDP_F3_ColorMatrix":
        arr = raw.reshape(V, count, 3, 3, 2)
    elif dtype == "QDP_F3_ColorVector":
        arr = raw.reshape(V, count, 3, 2)[:, 0] if count == 1 else raw.reshape(V, count, 3, 2)
    else:
        raise ValueError(dtype)
    return dims, arr


FILES = {
    "lat.sample.l8": "lat_l8.npy",
    "fat.l8": "fat_l8.npy",
    "ksvector.src.sample.l8": "ksvector_src_l8.npy",
    "ksvector.ans.MdaggerMinv.sample.l8": "ksvector_ans_MdaggerMinv_l8.npy",
    "ksvector.ans.Minv.sample.l8": "ksvector_ans_Minv_l8.npy",
    "colormatrix.mom.sample.l8": "colormatrix_mom_l8.npy",
}


# lattice files of MILC's own tests, kept byte for byte for the file-format
# readers: v5 in both byte orders, SciDAC, ILDG, a foreign ILDG file without
# checksum, a v5 file whose stored checksum does not match its data, and a
# SciDAC colour-vector file
IO_FILES = [
    "lat.sample.l4444", "lat.sample.l4444.scidac", "lat.sample.l4444.ildg",
    "lat.sample.l4448", "lat.sample.l4448.scidac", "lat.sample.l4448.foreign",
    "lat.sample.5p7", "lat.sample.l6666", "lat.sample.l6666.hisq", "ksvector.src.sample.l8",
    "lat.sample.l6666b650m010m050", "lat.sample.hisq.f211b670.l6666",
]


def main(ref="/root/reference"):
    for src, dst in FILES.items():
        dims, arr = scidac_field(os.path.join(ref, "binary_samples", src))
        assert dims == [8, 8, 8, 8]
        np.save(os.path.join(HERE, dst), arr)
        print(dst, arr.shape, arr.dtype)
    # MILC's rational functions for its HISQ and asqtad RHMC tests (MD / GR / FA per pseudofermion)
    for name in ("rationals.sample.su3_rhmc_hisq", "rationals.sample.su3_omelyan_rhmc",
                 "rationals.sample.su3_leapfrog", "rationals.sample.su3_rhmc_hisq_f211",
                 "rationals.sample.su3_rhmc_hisq_f2111"):
        rat = open(os.path.join(ref, "ks_imp_rhmc", "test", name)).read()
        open(os.path.join(HERE, name), "w").write(rat)
    os.makedirs(os.path.join(HERE, "io"), exist_ok=True)
    for name in IO_FILES:
        data = open(os.path.join(ref, "binary_samples", name), "rb").read()
        open(os.path.join(HERE, "io", name), "wb").write(data)
        print("io/" + name, len(data))


if __name__ == "__main__":
    main(*sys.argv[1:])

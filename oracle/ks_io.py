"""Oracle (test infrastructure only) for MILC's lattice-file formats: a numpy
restatement of the MILC version-5 binary gauge format and of the SciDAC / ILDG
(LIME) formats, with their checksums.  Only tests/ and bench.py's cpu_baseline
leg use it, as the checker of the library's native readers and writers.

Restates:
  MILC v5 gauge file      generic/io_lat_utils.c:370-389 (header: magic 20103,
                          dims[4], time_stamp[64], order = 88 bytes), then the
                          site list when order != 0, then the checksum (sum29,
                          sum31; io_lat_utils.c:766-812, io_lat4.c:570-585),
                          then 4 x fsu3_matrix per site; byte order from the
                          magic number, io_lat_utils.c:1305-1345
  v5 checksum             generic/io_lat4.c:694-707: every 32-bit data word, in
                          native order after byte reversal, rotated left by its
                          word rank mod 29 (sum29) and mod 31 (sum31) and XORed
  LIME records            include/file_types.h:10 (magic 0x456789ab); 144-byte
                          big-endian header (magic, version, flags, length,
                          128-byte type), payload padded to 8 bytes
  SciDAC / ILDG payload   big-endian, sites in lexicographic order (x fastest),
                          datacount objects per site (scidac-private-record-xml)
  SciDAC checksum         QIO's DML checksum (QIO is an external library, not in
                          the reference tree; its published algorithm): per site,
                          c = CRC-32 (IEEE, zlib's crc32) of the site's bytes as
                          stored, suma ^= rotl(c, rank % 29), sumb ^= rotl(c,
                          rank % 31), rank = lexicographic site index
  plaquette, link trace   generic/d_plaq4.c (average Re Tr of the space-space
                          and space-time plaquettes), generic/d_linktrsum.c
                          (printed as Re Tr / 3) -- used to pin site and
                          direction order against MILC's printed CHECK values

Pinned by: the checksums stored in the reference's binary_samples files, the
checksums and plaquettes MILC printed when it read them (*.sample-out), and
equality of the same configuration stored in v5, SciDAC and ILDG form.
"""
import re
import struct
import zlib

import numpy as np

V5_MAGIC = 20103
LIME_MAGIC = 0x456789AB


def rotl32(w, r):
    """rotate uint32 array w left by r (array), r in [0, 32)"""
    w = w.astype(np.uint64)
    r = r.astype(np.uint64)
    return (((w << r) | (w >> ((32 - r) % 32))) & 0xFFFFFFFF).astype(np.uint32)


def v5_checksum(words_native):
    """(sum29, sum31) of a uint32 array in native order (io_lat4.c:694-707)"""
    n = np.arange(words_native.size, dtype=np.uint64)
    s29 = np.bitwise_xor.reduce(rotl32(words_native, n % 29)) if words_native.size else np.uint32(0)
    s31 = np.bitwise_xor.reduce(rotl32(words_native, n % 31)) if words_native.size else np.uint32(0)
    return int(s29), int(s31)


def scidac_checksum(payload, nsites):
    """(suma, sumb) of a SciDAC binary payload with nsites equal-sized sites"""
    per = len(payload) // nsites
    crc = np.array([zlib.crc32(payload[i * per:(i + 1) * per]) for i in range(nsites)], dtype=np.uint32)
    rank = np.arange(nsites, dtype=np.uint64)
    return int(np.bitwise_xor.reduce(rotl32(crc, rank % 29))), int(np.bitwise_xor.reduce(rotl32(crc, rank % 31)))


def lime_records(data):
    out, off = [], 0
    while off + 144 <= len(data):
        magic, _ver, _flags, length = struct.unpack(">IHHQ", data[off:off + 16])
        if magic != LIME_MAGIC:
            raise ValueError(f"bad LIME magic at byte {off}")
        typ = data[off + 16:off + 144].split(b"\0")[0].decode()
        out.append((typ, data[off + 144:off + 144 + length]))
        off += 144 + ((length + 7) // 8) * 8
    return out


def read_v5(path):
    """MILC v5 gauge file -> dict(dims, links[V][4][3][3][2] float32 lexicographic,
    sum29, sum31 (stored), calc29, calc31, byte_reversed)"""
    data = open(path, "rb").read()
    m_le = struct.unpack("<i", data[:4])[0]
    if m_le == V5_MAGIC:
        e = "<"
    elif struct.unpack(">i", data[:4])[0] == V5_MAGIC:
        e = ">"
    else:
        raise ValueError("not a MILC v5 gauge file")
    dims = list(struct.unpack(e + "4i", data[4:20]))
    order = struct.unpack(e + "i", data[84:88])[0]
    V = int(np.prod(dims))
    off = 88
    sites = np.arange(V)
    if order != 0:
        sites = np.frombuffer(data[off:off + 4 * V], dtype=e + "i4").astype(np.int64)
        off += 4 * V
    s29, s31 = struct.unpack(e + "2I", data[off:off + 8])
    off += 8
    nword = V * 4 * 18
    words = np.frombuffer(data[off:off + 4 * nword], dtype=e + "u4").astype(np.uint32)
    c29, c31 = v5_checksum(words)
    vals = words.view(np.float32).reshape(V, 4, 3, 3, 2)
    links = np.empty_like(vals)
    links[sites] = vals
    return {"dims": dims, "links": links, "sum29": s29, "sum31": s31, "calc29": c29, "calc31": c31,
            "byte_reversed": e == ">"}


def _xml(rec, tag):
    m = re.search(rb"<" + tag.encode() + rb">([^<]*)</" + tag.encode() + rb">", rec)
    return m.group(1).decode().strip() if m else None


def read_lime_field(path):
    """SciDAC or ILDG lattice field -> dict(dims, data [V][count][...] float64 lexicographic,
    datatype, precision, suma/sumb stored (None if absent), calca/calcb)"""
    recs = lime_records(open(path, "rb").read())
    meta = {}
    for typ, payload in recs:
        meta.setdefault(typ, payload)
    if "ildg-binary-data" in meta:
        fmt = meta["ildg-format"]
        dims = [int(_xml(fmt, t)) for t in ("lx", "ly", "lz", "lt")]
        prec = "F" if int(_xml(fmt, "precision")) == 32 else "D"
        datatype, count, payload = "ColorMatrix", 4, meta["ildg-binary-data"]
    else:
        dims = [int(x) for x in _xml(meta["scidac-private-file-xml"], "dims").split()]
        rx = meta["scidac-private-record-xml"]
        prec, count = _xml(rx, "precision"), int(_xml(rx, "datacount"))
        datatype = _xml(rx, "datatype").split("_")[-1]
        payload = meta["scidac-binary-data"]
    V = int(np.prod(dims))
    raw = np.frombuffer(payload, dtype=">f4" if prec == "F" else ">f8").astype(np.float64)
    shape = (V, count, 3, 3, 2) if datatype == "ColorMatrix" else (V, count, 3, 2)
    out = {"dims": dims, "data": raw.reshape(shape), "datatype": datatype, "precision": prec,
           "suma": None, "sumb": None}
    if "scidac-checksum" in meta:
        out["suma"] = int(_xml(meta["scidac-checksum"], "suma"), 16)
        out["sumb"] = int(_xml(meta["scidac-checksum"], "sumb"), 16)
    out["calca"], out["calcb"] = scidac_checksum(payload, V)
    return out


def lex_coords(dims):
    V = int(np.prod(dims))
    l = np.arange(V)
    x = l % dims[0]
    y = (l // dims[0]) % dims[1]
    z = (l // (dims[0] * dims[1])) % dims[2]
    t = l // (dims[0] * dims[1] * dims[2])
    return x, y, z, t


def node_index_of_lex(dims):
    """MILC node_index of each lexicographic site (single rank, even sites first,
    generic/layout_hyper_prime.c:509-520)"""
    V = int(np.prod(dims))
    x, y, z, t = lex_coords(dims)
    l = np.arange(V)
    return np.where((x + y + z + t) % 2 == 0, l // 2, (l + V) // 2)


def plaquette(links, dims):
    """(ss, st): average Re Tr over space-space and space-time plaquettes of
    links[V][4][3][3][2] in lexicographic order (generic/d_plaq4.c)"""
    U = links[..., 0] + 1j * links[..., 1]
    X = np.asarray(dims)
    x, y, z, t = lex_coords(dims)
    c = np.stack([x, y, z, t], axis=1)

    def shift(mu):
        d = c.copy()
        d[:, mu] = (d[:, mu] + 1) % X[mu]
        return d[:, 0] + X[0] * (d[:, 1] + X[1] * (d[:, 2] + X[2] * d[:, 3]))

    ss, st, nss, nst = 0.0, 0.0, 0, 0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            Um, Un = U[:, mu], U[:, nu]
            p = Um @ Un[shift(mu)] @ np.conj(np.swapaxes(Um[shift(nu)], 1, 2)) @ np.conj(np.swapaxes(Un, 1, 2))
            tr = float(np.sum(np.trace(p, axis1=1, axis2=2).real))
            if nu == 3:
                st, nst = st + tr, nst + 1
            else:
                ss, nss = ss + tr, nss + 1
    V = int(np.prod(dims))
    return ss / (nss * V), st / (nst * V)


def link_trace(links):
    """average Re Tr(U) / 3 over all links (generic/d_linktrsum.c, as printed)"""
    return float(np.mean(links[..., 0, 0, 0] + links[..., 1, 1, 0] + links[..., 2, 2, 0])) / 3.0

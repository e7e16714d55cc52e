"""Lattice files through the C-ABI of include/milc_io.h (libmilc_hip.so; host
code, no GPU needed), with MILC's names and argument meaning:

* ``reload_lattice(filename, dims)``  -- generic/io_helpers.c:145 (the file
  readers behind it: restore_serial for v5, restore_serial_scidac for SciDAC
  and ILDG); returns the links in MILC's layout, node_index order,
  (V, 4, 3, 3, 2) float64, without KS phases, and the file info.
* ``save_lattice(flag, filename, links, dims)``  -- generic/io_helpers.c:20
  with SAVE_SERIAL (v5), SAVE_SERIAL_SCIDAC or SAVE_SERIAL_ILDG.
* ``restore_ks_vector_scidac_to_field`` / ``save_ks_vector_scidac_from_field``
  -- generic_ks/io_scidac_ks.c:166 / :67, (V, 3, 2) float64.
* ``restore_color_matrix_scidac_to_field`` /
  ``save_color_matrix_scidac_from_field`` -- generic/io_scidac.c:572 / :683.

Errors raise MilcHipError with the library's status code.
"""
import ctypes

import numpy as np

from . import ks

SAVE_SERIAL = 1          # MILC v5 binary (always 32-bit)
SAVE_SERIAL_SCIDAC = 2
SAVE_SERIAL_ILDG = 3

FORMAT_NAMES = {1: "v5", 2: "scidac", 3: "ildg"}


class IoInfo(ctypes.Structure):
    _fields_ = [("format", ctypes.c_int), ("precision", ctypes.c_int), ("dims", ctypes.c_int * 4),
                ("count", ctypes.c_int), ("datatype", ctypes.c_int), ("byte_reversed", ctypes.c_int),
                ("has_checksum", ctypes.c_int), ("checksum_ok", ctypes.c_int),
                ("stored_a", ctypes.c_uint), ("stored_b", ctypes.c_uint),
                ("computed_a", ctypes.c_uint), ("computed_b", ctypes.c_uint)]

    def as_dict(self):
        return {"format": FORMAT_NAMES.get(self.format, self.format), "precision": self.precision,
                "dims": list(self.dims), "count": self.count,
                "datatype": {1: "ColorMatrix", 2: "ColorVector"}.get(self.datatype, self.datatype),
                "byte_reversed": bool(self.byte_reversed), "has_checksum": bool(self.has_checksum),
                "checksum_ok": bool(self.checksum_ok), "stored": (self.stored_a, self.stored_b),
                "computed": (self.computed_a, self.computed_b)}


_bound = None


def _lib():
    global _bound
    L = ks.load_library()
    if _bound is not L:
        P, I = ctypes.c_void_p, ctypes.c_int
        S, PI, PInfo = ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(IoInfo)
        L.milc_io_read_info.argtypes = [S, PInfo]
        L.milc_io_read_gauge.argtypes = [S, PI, P, PInfo]
        L.milc_io_write_gauge.argtypes = [S, I, I, PI, P, PInfo]
        L.milc_io_read_ksvector.argtypes = [S, PI, P, PInfo]
        L.milc_io_write_ksvector.argtypes = [S, I, PI, P, PInfo]
        L.milc_io_read_colormatrix.argtypes = [S, PI, I, P, PInfo]
        L.milc_io_write_colormatrix.argtypes = [S, I, PI, I, P, PInfo]
        _bound = L
    return L


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(d) for d in dims])


def _path(p):
    return str(p).encode()


def _buf(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def read_info(filename):
    """format, dimensions and checksums of a lattice file"""
    info = IoInfo()
    ks._check(_lib().milc_io_read_info(_path(filename), ctypes.byref(info)), "milc_io_read_info")
    return info.as_dict()


def reload_lattice(filename, dims=None):
    """(links (V, 4, 3, 3, 2) node_index order, info); dims default to the file's"""
    if dims is None:
        dims = read_info(filename)["dims"]
    V = int(np.prod(dims))
    links = np.zeros((V, 4, 3, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_read_gauge(_path(filename), _dims(dims), links.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(info)), "milc_io_read_gauge")
    return links, info.as_dict()


def save_lattice(flag, filename, links, dims, precision=1):
    """flag: SAVE_SERIAL (v5), SAVE_SERIAL_SCIDAC or SAVE_SERIAL_ILDG; returns the info written"""
    V = int(np.prod(dims))
    links = _buf(links, (V, 4, 3, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_write_gauge(_path(filename), flag, precision, _dims(dims),
                                         links.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info)),
              "milc_io_write_gauge")
    return info.as_dict()


def restore_ks_vector_scidac_to_field(filename, dims):
    V = int(np.prod(dims))
    v = np.zeros((V, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_read_ksvector(_path(filename), _dims(dims), v.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.byref(info)), "milc_io_read_ksvector")
    return v, info.as_dict()


def save_ks_vector_scidac_from_field(filename, vec, dims, precision=1):
    V = int(np.prod(dims))
    vec = _buf(vec, (V, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_write_ksvector(_path(filename), precision, _dims(dims),
                                            vec.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info)),
              "milc_io_write_ksvector")
    return info.as_dict()


def restore_color_matrix_scidac_to_field(filename, dims, count):
    V = int(np.prod(dims))
    m = np.zeros((V, count, 3, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_read_colormatrix(_path(filename), _dims(dims), count,
                                              m.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info)),
              "milc_io_read_colormatrix")
    return m, info.as_dict()


def save_color_matrix_scidac_from_field(filename, mat, dims, count, precision=1):
    V = int(np.prod(dims))
    mat = _buf(mat, (V, count, 3, 3, 2))
    info = IoInfo()
    ks._check(_lib().milc_io_write_colormatrix(_path(filename), precision, _dims(dims), count,
                                               mat.ctypes.data_as(ctypes.c_void_p), ctypes.byref(info)),
              "milc_io_write_colormatrix")
    return info.as_dict()

/* milc_io.h -- C ABI of the lattice-file readers and writers: MILC's version-5
 * binary gauge format and the SciDAC / ILDG (LIME) formats, for gauge links,
 * colour-vector fields and colour-matrix fields (momenta, forces).  Host code
 * (file I/O), part of libmilc_hip.so; no GPU is needed.
 *
 * Fields cross the ABI in MILC's own in-memory layout: double precision,
 * sites in node_index order (single rank: lexicographic index l = x + nx(y +
 * ny(z + nz t)), even sites first, i = l/2 even, (l+V)/2 odd;
 * generic/layout_hyper_prime.c:509-520), gauge links as su3_matrix[4*i+dir],
 * no KS phases (files hold the bare links).
 *
 * Return values: 0 success; -1 the file cannot be opened, read or written, or
 * is truncated; -2 unknown format or unsupported content; -3 checksum
 * mismatch (SciDAC / ILDG: the data are not returned as valid, as QIO fails
 * the read); -4 the file's dimensions differ from dims; -5 bad argument.
 * A v5 checksum mismatch is reported in info->checksum_ok = 0 with return 0:
 * MILC prints the violation and continues (generic/io_lat_utils.c:780-788). */
#ifndef MILC_IO_H
#define MILC_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define MILC_IO_V5 1
#define MILC_IO_SCIDAC 2
#define MILC_IO_ILDG 3

typedef struct {
  int format;         /* MILC_IO_V5 / _SCIDAC / _ILDG */
  int precision;      /* 1: 32-bit payload, 2: 64-bit */
  int dims[4];
  int count;          /* objects per site (4 for gauge links) */
  int datatype;       /* 1: colour matrix, 2: colour vector */
  int byte_reversed;  /* v5: file written in the other byte order */
  int has_checksum;
  int checksum_ok;
  unsigned int stored_a, stored_b;    /* v5: sum29, sum31; LIME: suma, sumb */
  unsigned int computed_a, computed_b;
} milc_io_info;

/* Header, format, dimensions and checksums (the checksums need a full pass
 * over the payload) of a gauge, vector or matrix file. */
int milc_io_read_info(const char* path, milc_io_info* info);

/* Gauge configuration, any of the three formats (format detected from the
 * magic number).  Replaces the file readers behind reload_lattice
 * (generic/io_helpers.c:145): restore_serial (generic/io_lat4.c:1820, v5) and
 * restore_serial_scidac (generic/io_scidac.c:504 -> restore_scidac :365,
 * SciDAC and ILDG through QIO).  links: V*4*18 doubles. */
int milc_io_read_gauge(const char* path, const int dims[4], double* links, milc_io_info* info);

/* Writes a gauge configuration, as save_lattice (generic/io_helpers.c:20)
 * with SAVE_SERIAL -> save_serial (generic/io_lat4.c:1888; v5 is always
 * 32-bit), SAVE_SERIAL_SCIDAC -> save_serial_scidac (generic/io_scidac.c:308)
 * or SAVE_SERIAL_ILDG -> save_serial_ildg (:328); precision 1 or 2 for the
 * LIME formats.  info (nullable) receives the checksums written. */
int milc_io_write_gauge(const char* path, int format, int precision, const int dims[4], const double* links,
                        milc_io_info* info);

/* SciDAC colour-vector field (QDP_F3/D3_ColorVector, one per site), as
 * restore_ks_vector_scidac_to_field (generic_ks/io_scidac_ks.c:166) and
 * save_ks_vector_scidac_from_field (:67).  vec: V*6 doubles. */
int milc_io_read_ksvector(const char* path, const int dims[4], double* vec, milc_io_info* info);
int milc_io_write_ksvector(const char* path, int precision, const int dims[4], const double* vec,
                           milc_io_info* info);

/* SciDAC colour-matrix field with `count` matrices per site (momenta and
 * forces, e.g. check_fermion_force's output), as
 * restore_color_matrix_scidac_to_field (generic/io_scidac.c:572) and
 * save_color_matrix_scidac_from_field (:683).  mat: V*count*18 doubles. */
int milc_io_read_colormatrix(const char* path, const int dims[4], int count, double* mat, milc_io_info* info);
int milc_io_write_colormatrix(const char* path, int precision, const int dims[4], int count, const double* mat,
                              milc_io_info* info);

#ifdef __cplusplus
}
#endif

#endif

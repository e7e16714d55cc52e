/* milc_hip_glue.h -- shared by the MILC glue files of this directory.
 *
 * The glue files are added to a MILC 7.8.1 tree in place of the QUDA glue
 * objects (generic_ks/d_congrad5_fn_gpu.c, generic_ks/ks_multicg_offset_gpu.c
 * and the GPU wrapper of generic_ks/fermion_force_hisq_multi.c) and compiled
 * with MILC's own headers; MILC's callers reach them through the renaming
 * macros of include/imp_ferm_links.h:73-77,238-239 (-DUSE_CG_GPU).
 */
#ifndef MILC_HIP_GLUE_H
#define MILC_HIP_GLUE_H

/* Sets up the library's lattice from MILC's globals (nx, ny, nz, nt) on the
 * first call: the role initialize_quda() plays for the QUDA glue
 * (generic/milc_to_quda_utilities.c:11-43).  The library runs one lattice on
 * one GPU, so a run on more than one rank stops here.  Returns 0 or stops
 * the program with terminate(). */
int milc_hip_glue_init(void);

/* Stops the program with terminate(1) after printing `what` and the library's
 * return code, as MILC's glue does on a failed offload. */
void milc_hip_glue_fail(const char* what, int rc);

/* Sites [lo, hi) of one parity in MILC's site order (even sites first);
 * EVENANDODD gives every site. */
void milc_hip_glue_parity_range(int parity, int* lo, int* hi);

#endif

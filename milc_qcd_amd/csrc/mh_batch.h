// mh_batch.h -- many right-hand sides at once: batched colour-vector fields,
// the batched fat + Naik stencil and the batched solvers built on it.
//
// Batched field of one parity, NR right-hand sides (NR = 4, 8 or 16):
//   record (site cb, rhs r) = 6 reals  [re c0, re c1, re c2, im c0, im c1, im c2]
//   at offset (cb * NR + r) * 6; R = float (F32) or double (F64);
//   16-bit (P_H16): one int4 per record at cb * NR + r (mh_bkern.h).
// All NR records of a site are contiguous, so a wavefront that spreads the
// right-hand sides of a few sites over its lanes reads whole cache lines, and
// the link matrices of a site are fetched once for all NR vectors.
#pragma once
#include <vector>

#include "mh_internal.h"

namespace mh {

#define BNR_MAX 16
#define BMS_MAX 32   // shifts per multi-shift solve
#define BKW 8        // iterations per deferred shift-update window

enum BKind { BK_STORE = 0, BK_MS = 1, BK_CG = 2 };

// One application of the batched stencil on output parity pb:
//   dst = D src                         (BK_STORE)
//   dst = D src + coef[r] p, <p,dst>    (BK_MS: coef = -offset of the base shift)
//   dst = D src + coef[r] p, <p,dst>, <r,dst>, <dst,dst>   (BK_CG: coef = -offset)
// Dots are real parts, summed per right-hand side into
// partials[(v * NR + r) * pstride + block].
struct BDOp {
  int prec = P_F32;
  int nr = 16;
  int kind = BK_STORE;
  int pb = 0;
  const void* src = nullptr;
  void* dst = nullptr;
  const void* p = nullptr;
  const void* r = nullptr;
  double coef[BNR_MAX] = {};
  double* partials = nullptr;
  const int* halt_all = nullptr;  // kernels return at once when *halt_all != 0
  int variant = -1;               // -1: the autotuned choice
};

// A stencil variant: VALU (tps threads per site, nr/tps right-hand sides per
// thread) or MFMA (v_mfma_f32_4x4x1f32, loop sites groups per wave).
struct BVariant {
  int mfma, tps, block, loop, order;  // order: site-order block shape index (0 = checkerboard order)
};

size_t bvec_bytes(int prec, int nr);
int balloc(void** p, int prec, int nr);          // zeroed batched field
int bdslash(const BDOp& op, int* nblocks_out);  // enqueue on the context stream
int bpartials(double** p, int* pstride);         // per-rhs partial-sum buffer
// sums[ch] = sum over the nblocks partial records of channel ch (ch < nch)
int breduce(int nch, int nblocks, double* sums);
std::vector<BVariant> bvariants(int prec, int nr);
int btune(int prec, int nr, int* chosen, double* usec_best);  // autotune (timed with HIP events)
int btime_variant(int prec, int nr, int variant, int reps, double* usec);
void bdescribe(int prec, int nr, int variant, char* buf, int n);
void batch_release();

// pack / unpack between single-rhs FP64 fields and batched fields
int bpack(const VecField& in, void* batched, int prec, int nr, int r);
int bunpack(const void* batched, int prec, int nr, int r, VecField& out);

// ---- solvers ------------------------------------------------------------------
// Sloppy precision of the batched mixed CG (per-shift refinement after a
// batched multi-shift, block solves): 16-bit records and links.  24^3x48, 16
// sources x 12 shifts refined from the FP32 stage to 1e-9: same iteration
// counts as FP32 sloppy, 0.347-0.351 s of refinement against 0.426-0.432 s.
constexpr int kBatchSloppy = P_H16;

struct BMSParams {
  int parity;          // MILC code of the solve parity
  int max_iters;
  double resid;        // base-system |r|/|b| target of the multi-shift stage
  double refine_resid; // true |r|/|b| target per shift after refinement (0: none)
  double delta;        // reliable-update threshold of the refinement
  int refine_sloppy = kBatchSloppy;  // sloppy precision of the refinement CG (BCGParams::sloppy)
};
struct BMSStats {
  int ms_iters = 0;                 // multi-shift iterations (max over rhs)
  int refine_iters[BMS_MAX] = {};   // refinement iterations per shift (max over rhs)
  double true_rel[BMS_MAX] = {};    // max over rhs of |b - A_j x_j| / |b| per shift
  double true_rel_r[BMS_MAX][BNR_MAX] = {};  // per shift and rhs
  double ms_seconds = 0, refine_seconds = 0;
  int converged = 0;
};
// x[j] (batched FP64, nr rhs) = (-D_eo D_oe + offsets[j])^-1 b for every rhs.
int bmultishift(const void* b64, int nr, const std::vector<double>& offsets, const BMSParams& prm,
                std::vector<void*>& x64, BMSStats& st);
// The same with a shift set per rhs: x[j] rhs r = (-D_eo D_oe + offs[j nr + r])^-1 b_r,
// the FP32 stage of rhs r to stage_resid[r], then every shift refined to
// refine_resid[r] (nullptr: no refinement).  prm.resid / refine_resid unused.
int bmultishift_sets(const void* b64, int nr, int n, const double* offs, const double* stage_resid,
                     const double* refine_resid, const BMSParams& prm, std::vector<void*>& x64, BMSStats& st);

struct BCGParams {
  int parity;
  int max_iters;
  double resid;       // true |r|/|b| target (all rhs)
  double delta;       // reliable-update threshold
  const double* resid_r = nullptr;  // per-rhs targets (host array of nr), overriding resid
  int sloppy = kBatchSloppy;  // sloppy operator and search direction: P_F32 or P_H16 (16-bit records)
};
struct BCGStats {
  int iters = 0;               // max over rhs
  int reliable_updates = 0;
  double true_rel[BNR_MAX] = {};
  int converged[BNR_MAX] = {};
  double seconds = 0;
};
// Mixed-precision batched CG: FP32 (or 16-bit operator and search direction:
// prm.sloppy) sloppy iterations, FP64 reliable updates; x64 holds the initial
// guess; offsets[r] = 4 m^2 (+ sigma) per rhs.
int bcg_mixed(const void* b64, void* x64, int nr, const double* offsets, const BCGParams& prm, BCGStats& st);

// Refinement of one source's multi-shift solutions x[j] (FP64 fields of the
// solve parity) to the FP64 true |r|/|b| targets resid[j]: groups of up to four
// consecutive shifts as one batched mixed CG with the sloppy precision
// `sloppy` (P_F32 or P_H16).  Adds the iterations (each group once) to *iters,
// the final true relative residuals to fres[j] (when given) and ANDs the
// convergence flags into *converged.
int refine_shifts(const VecField& b, std::vector<VecField>& x, const double* offs, const double* resid, int parity,
                  int max_iters, int sloppy, long* iters, double* fres, int* converged);

}  // namespace mh

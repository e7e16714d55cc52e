// mh_internal.h -- host-side declarations shared by the library's translation units.
#pragma once
#include <functional>
#include <initializer_list>
#include <vector>
#include "mh_common.h"

namespace mh {

// Device-resident scalars of one Krylov solve (all FP64).  The reduction
// kernels write them and the vector-update kernels read them, so the host
// never sits between two iterations.
struct CGScalars {
  double rsq;          // recurrence residual |r|^2 (FEWSUMS, d_congrad5_fn_milc.c:339)
  double actual_rsq;   // explicitly summed |r|^2 of the last update
  double oldrsq;
  double pkp, c_tr, c_tt;
  double a, b;
  double source_norm;
  double rsq_stop;     // stop when rsq < rsq_stop (absolute)
  double rsq_first;    // |r|^2 right after the last reliable update
  double rsq_max;      // max |r|^2 since the last reliable update
  double relrsq;       // Fermilab relative residual of the last update or
                       // restart: sqrt(sum_i |r_i|^2/|x_i|^2 / Vh), the value
                       // of my_relative_residue (d_congrad5_fn_milc.c:37-56)
  double relrsq_stop;  // with rel_mode: stop once relrsq < relrsq_stop too
  int rel_mode;        // 1: the halt test needs relrsq, taken after the update
  int halt;            // set by the scalar step once the target is met: later
                       // iterations' kernels return immediately
  int go;              // written by every scalar step: 1 if this iteration's
                       // vector update must run
  int iters;           // scalar steps executed since the last reset
};

// Reduction scratch: partial sums per block (FP64) and the scalar block.
struct Workspace {
  double* partials = nullptr;  // [max_blocks * 4]
  int max_blocks = 0;
  CGScalars* scal = nullptr;   // device
  CGScalars* scal_host = nullptr;  // pinned host mirror
  void* pinned = nullptr;          // pinned host scratch for small read-backs
  size_t pinned_bytes = 0;
};

// A set of links on the device in one precision.
struct DevLinks {
  LinkField fat{}, lng{};
  int prec = -1;
  bool valid = false;
};

// Plane stride (elements) of the per-parity fields: Vh rounded up to 64, plus
// 192 when that is a multiple of 256.  Lattices such as 32^3x64 (Vh = 2^20)
// otherwise put the planes of a field, and the fields of one pool, a large
// power of two bytes apart, so that the same cb of every plane falls in the
// same memory channel; the padded stride spreads them (16-bit Dslash at
// 32^3x64: 109.6 us with stride Vh, 99.5 us with Vh + 192).
inline int plane_stride(int Vh) {
  const int s = (Vh + 63) / 64 * 64;
  return s % 256 == 0 ? s + 192 : s;
}

struct Context {
  int device = 0;
  int dims[4] = {0, 0, 0, 0};
  Geom geom{};
  int Vh = 0;
  int stride = 0;     // padded plane stride (elements) for per-parity fields
  hipStream_t stream = nullptr;
  Workspace ws;
  // link cache (keyed like d_congrad5_fn_gpu.c:120-126: fn pointer + refresh flag)
  const void* link_key = nullptr;
  const void* link_key2 = nullptr;
  DevLinks links[3];   // by Prec
  unsigned links_version = 0;  // bumped whenever the FP64 links change
  LinkField thin{};            // thin links (phases in) when the device built the links
  bool thin_valid = false;
  double* staging = nullptr;   // FP64 AoS staging buffer, 4*V*18 doubles
  size_t staging_elems = 0;
  int block = 256;
  bool initialised = false;
};

Context& ctx();

// A hipGraph captured from the kernels that `enqueue` puts on the context
// stream (no allocations, memsets or synchronisations inside), replayed with
// hipGraphLaunch.
struct HipGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  HipGraph() = default;
  HipGraph(const HipGraph&) = delete;
  HipGraph& operator=(const HipGraph&) = delete;
  ~HipGraph();
};
int capture_graph(HipGraph& gr, const std::function<int()>& enqueue);

// device memory helpers
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);

// Scratch of the allocation-free solve: a static device array of
// kScratchBytes per device.  scratch_init looks up its address on `device`
// (the current device) once; while a ScratchArena is alive, alloc_vec carves
// zeroed fields from that array, dev_alloc refuses (-7) and the pinned host
// scratch does not grow.
constexpr size_t kScratchBytes = (size_t)256 << 20;
int scratch_init(int device);
struct ScratchArena {
  bool ok = false;  // false: no scratch address for the current device
  ScratchArena();
  ~ScratchArena();
  ScratchArena(const ScratchArena&) = delete;
  ScratchArena& operator=(const ScratchArena&) = delete;
};
// Work-vector bytes cg_solve takes from the arena for sloppy precision ps.
// per-kernel times of the 16-bit mixed CG iteration in its own sequence (us[5])
int cg_iteration_profile(const VecField& src, double mass, int parity, int reps, double us[5]);
size_t cg_scratch_bytes(int ps, int stride);
// hand the pooled free blocks back to the device (end of context)
void dev_pool_release();
int alloc_vec(VecField& v, int prec, int stride);
void free_vec(VecField& v);
// Frees the vector fields it holds when it goes out of scope, so that every
// return path of a solver hands its blocks back to the pool.
struct VecScope {
  std::vector<VecField*> v;
  VecScope() = default;
  VecScope(std::initializer_list<VecField*> l) : v(l) {}
  VecScope(const VecScope&) = delete;
  VecScope& operator=(const VecScope&) = delete;
  void add(VecField* f) { v.push_back(f); }
  ~VecScope() {
    for (VecField* f : v) free_vec(*f);
  }
};
// Runs a clean-up action when it goes out of scope (every return path).
template <class F>
struct OnExit {
  F f;
  explicit OnExit(F fn) : f(fn) {}
  OnExit(const OnExit&) = delete;
  OnExit& operator=(const OnExit&) = delete;
  ~OnExit() { f(); }
};
// Pinned host scratch of at least `bytes` (grown on demand, freed with the
// context): small device-to-host read-backs land in pinned memory, so the
// runtime copies them by DMA and never through a pageable bounce.
void* pinned_scratch(size_t bytes);
int alloc_links(LinkField& l, int prec, int stride);
void free_links(LinkField& l);
size_t vec_bytes(int prec, int stride);
size_t link_bytes(int prec, int stride);

// ---- layout ----------------------------------------------------------------
int upload_vec_milc(const double* host_aos, VecField& out, int prec, int pb);
int download_vec_milc(const VecField& in, int prec, int pb, double* host_aos);
int vec_from_milc_dev(const double* dev_aos, VecField& out, int prec, int pb);
int vec_to_milc_dev(const VecField& in, int prec, int pb, double* dev_aos);
int links_from_milc_dev(const double* dev_aos, LinkField& out, int prec);
int links_to_milc_dev(const LinkField& in, int prec, double* dev_aos);
int convert_links(const LinkField& in, int pin, LinkField& out, int pout);
int convert_long_h16r(const LinkField& in, LinkField& out, double tol);
int convert_vec(const VecField& in, int pin, VecField& out, int pout);

// ---- stencil ---------------------------------------------------------------
// dst(pb_out) = D src(1-pb_out)
int dslash(const DevLinks& L, const VecField& src, VecField& dst, int pb_out, int prec_vec);

// ---- solver entry points (device fields, FP64 x/src) -------------------------
struct SolveParams {
  int parity;          // MILC code
  int max_iters;       // per restart (qic->max)
  int max_restarts;    // qic->nrestart
  double resid;        // target |r|/|b| (0: no test on it)
  double relresid;     // Fermilab relative residual target (qic->relresid, 0:
                       // no test); met when my_relative_residue < relresid^2,
                       // the test of d_congrad5_fn_milc.c:92,213-220
  int sloppy_prec;     // P_F64 (plain), P_F32 or P_H16
  double delta;        // reliable-update threshold (mixed)
  int batch;           // iterations enqueued per host check
};
// Reliable-update threshold of the mixed CG: each sloppy phase runs until its
// iterated residual has fallen by this factor, then x and r are rebuilt in
// FP64.  A 16-bit sloppy residual keeps about 3e-5 relative resolution, so a
// 3e3 drop leaves it accurate to about ten percent; the stopping test is the
// FP64 true residual either way.  (32^3x64, m = 0.01, 1e-10, measured on one
// box: 1e-3 gives 365-368 iterations and 93.0-94.4 ms per solve, 3e-4
// 359-365 iterations and 91.4 ms, 1e-4 368-375 iterations and 96.3 ms.)
constexpr double kReliableDelta = 3e-4;
// The same threshold for the batched mixed CG, whose sloppy phase runs in
// FP32 (24^3x48, 16 sources x 12 shifts refined to 1e-9: 0.1 takes 0.471 s of
// refinement, 1e-2 0.438 s, 1e-3 0.433 s).
constexpr double kBatchReliableDelta = 1e-3;
// Base-system residual of an FP32 multi-shift stage that per-shift refinement
// takes to an FP64 target below it: the refinement starts from the stage
// solutions, and a deeper stage is cheaper than more refinement (same
// workload: stage 1e-7 with 176 multi-shift iterations, 84 refinement
// iterations for the lightest shift, 0.749 s; 1e-8 with 214 and 58, 0.706 s;
// 3e-9 stalls at the FP32 floor: 219 and 57).  One source, 48^3x96, 12
// shifts, refined in groups of four to 1e-10: 1.92 s with the stage at 1e-7,
// 1.85 s at 1e-8.
constexpr double kFp32StageResid = 1e-8;

struct SolveStats {
  double final_rsq = 0, final_relrsq = 0, size_r = 0;
  int iters = 0, restarts = 0, converged = 0, reliable_updates = 0;
  double seconds = 0;
};
int cg_solve(const VecField& src, VecField& x, double mass, const SolveParams& sp, SolveStats& st);
// builds the sloppy copy of the FP64 device links in precision prec if absent
int ensure_links(int prec);

struct MultiParams {
  int parity;
  int max_iters;       // total iteration cap (qic->max * qic->nrestart)
  double resid;        // base-system target |r|/|b|
  int prec;            // P_F64 or P_F32 bulk
  int batch;
};
int multicg_solve(const VecField& src, std::vector<VecField>& x, const std::vector<double>& offsets,
                  const MultiParams& mp, SolveStats& st);

// ---- force outer products ---------------------------------------------------
int force_oprod(const std::vector<VecField>& xe, const std::vector<VecField>& xo,
                const std::vector<double>& residues, LinkField& o1, LinkField& o3);
// HISQ path coefficients (generic_ks/imp_actions/hisq/hisq_u3_action.h) in the
// order {one_link, naik, 3-staple, 5-staple, 7-staple, Lepage}
constexpr double kHisqLevel2[6] = {(1.0 / 8.0) + (2.0 * 6.0 / 16.0) + (1.0 / 8.0), -1.0 / 24.0, -1.0 / 8.0 * 0.5,
                                   1.0 / 8.0 * 0.25 * 0.5, -1.0 / 8.0 * 0.125 * (1.0 / 6.0), -2.0 / 16.0};
constexpr double kHisqLevel1[6] = {1.0 / 8.0, 0.0, -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5,
                                   -1.0 / 8.0 * 0.125 * (1.0 / 6.0), 0.0};
constexpr double kForceFilter = 5.0e-5;  // HISQ_FORCE_FILTER of ks_imp_rhmc/Make_template
int hisq_force_from_oprod(const LinkField& U, const LinkField& V, const LinkField& W, const LinkField& o1,
                          const LinkField& o3, const double coeff2[6], const double coeff1[6], double scale,
                          double filter, double* mom_dev, const LinkField* one_link_extra = nullptr);
// y = b y + a x on FP64 link fields (x negated on odd sites when signed_odd)
int link_axpy_dev(LinkField& y, const LinkField& x, double a, double b, int signed_odd);
// "1-link + Naik" table of a second Naik set (imp_actions/hisq/hisq_u3_action.h:86-101)
constexpr double kHisqNaikSet[2] = {1.0 / 8.0, -1.0 / 24.0};
// thin-link force adjoint form: 1 staple form (default), 0 path-table gather
void set_force_form(int form);
// One RHMC trajectory on the device links (one HISQ pseudofermion): rational
// functions in MILC's layout, res[0] = constant, res[1..order], pole[1..order]
struct RhmcRat {
  int order;
  double res[33], pole[33];
};
struct RhmcPseudo {
  double mass;  // 4 mass^2 is added to every pole
  RhmcRat md, gr, fa;
  double resid_md, resid_gr, resid_fa;
  double naik_eps;  // naik_term_epsilon: links X_2 + naik_eps X_3
};
#define RHMC_MAX_PSEUDO 8
struct RhmcParams {
  int steps;
  int integrator;  // MILC's int_alg_t: 0 leapfrog, 1 Omelyan, 4 2G1F, 5 3G1F, 6 5G1F, 7 6G1F
  double epsilon, beta, loop_coeff[3];
  int action;  // 0 HISQ, 1 asqtad (tadpole factor u0)
  double u0;
  int n_pseudo;
  RhmcPseudo pf[RHMC_MAX_PSEUDO];
  int mixed_md, mixed_action, action_from_md, accept, site_rng;
  uint64_t seed;
};
// stats: 32 doubles (see milc_hip_dev_rhmc_trajectory)
int rhmc_trajectory(const RhmcParams& p, double* stats);
// the next trajectory fails with -90 at this stage (1 heatbath, 2 start actions, 3 molecular dynamics)
void rhmc_fail_at(int stage);

// MILC's per-site generators (SITERAND) on the device
int prn_seed_dev(int iseed, int single);
bool prn_valid();
int ranmom_prn_dev(double* mom_dev);
int gaussian_prn_dev(int pb, VecField& v, double* sum);
double prn_node_uniform();
void prn_release();
// FN (asqtad) thin-link force from the level-0 outer products: mom = TA(scale U F)
int fn_force_from_oprod(const LinkField& U, const LinkField& o1, const LinkField& o3, const double coeff[6],
                        double scale, double* mom_dev);
// reunit_su3 on every link (KS phases taken out and put back when phases_in)
int reunitarize_dev(LinkField& U, int phases_in, double tol, int* errors, double* maxdev);

// Symanzik-improved gauge action on links without KS phases (coeff: plaquette,
// rectangle, parallelogram), its force into packed momenta, the link update and
// the momentum action
int gauge_force_dev(const LinkField& U, const double coeff[3], double eb3, double* mom_dev);
int gauge_action_dev(const LinkField& U, const double coeff[3], double* action);
int update_u_dev(LinkField& U, const double* mom_dev, double eps);
int mom_action_dev(const double* mom_dev, double* action);
void gauge_table_release();
// device path table of the FN force (freed with the context)
void fn_table_release();
// batched-solver caches (freed with the context)
void batch_release();

// ---- links ------------------------------------------------------------------
int gauge_random_dev(uint64_t seed, double scale, LinkField& U);
// V_out / W_out (nullable) receive the fat7 links and their U(3) projection
int hisq_links_dev(const LinkField& U, LinkField& fat, LinkField& lng, LinkField* W_out, int* nsvd,
                   LinkField* V_out = nullptr);
int asqtad_links_dev(const LinkField& U, double u0, LinkField& fat, LinkField& lng);
int hisq_aux_dev(const LinkField& U, LinkField& V, LinkField& W);
int rephase_dev(LinkField& U);
int gaussian_vec_dev(uint64_t seed, int pb, VecField& v);
// Gaussian anti-Hermitian momenta (packed, node_index order), weight exp(-|P|^2/2)
int ranmom_dev(uint64_t seed, double* mom);

// ---- BLAS --------------------------------------------------------------------
int vec_norm2(const VecField& v, int prec, double* out);  // synchronous
int vec_zero(VecField& v, int prec);
int vec_copy(const VecField& in, VecField& out, int prec);

}  // namespace mh

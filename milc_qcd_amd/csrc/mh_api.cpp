// mh_api.cpp -- extern "C" entry points (include/milc_hip.h).
//
// Each entry point restates the behaviour of the MILC GPU glue it replaces;
// the citations are in the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/milc_hip.h"
#include "mh_batch.h"
#include "mh_internal.h"

using namespace mh;

namespace {

const int kSlots = 64;
VecField g_slots[kSlots];

int parity_from_evenodd(int evenodd) {
  if (evenodd == MILC_HIP_EVEN_PARITY) return EVEN;
  if (evenodd == MILC_HIP_ODD_PARITY) return ODD;
  return -1;
}

// Upload a MILC host link pair (float or double) into the FP64 device links.
int upload_links(int ext_prec, const void* fat, const void* lng) {
  Context& c = ctx();
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  std::vector<double> conv;
  const double* hf = (const double*)fat;
  const double* hl = (const double*)lng;
  if (ext_prec == 1) {
    conv.resize(2 * n);
    for (size_t i = 0; i < n; i++) conv[i] = ((const float*)fat)[i];
    for (size_t i = 0; i < n; i++) conv[n + i] = ((const float*)lng)[i];
    hf = conv.data();
    hl = conv.data() + n;
  }
  for (int p = 0; p < 3; p++) {
    if (c.links[p].valid) {
      free_links(c.links[p].fat);
      free_links(c.links[p].lng);
      c.links[p].valid = false;
    }
  }
  if (c.thin_valid) {  // the host supplied fat/long links: no thin links on the device
    free_links(c.thin);
    c.thin_valid = false;
  }
  DevLinks& L = c.links[P_F64];
  int rc;
  if ((rc = alloc_links(L.fat, P_F64, c.stride)) || (rc = alloc_links(L.lng, P_F64, c.stride))) return rc;
  double* stage = nullptr;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  MH_CHECK(hipMemcpyAsync(stage, hf, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if ((rc = links_from_milc_dev(stage, L.fat, P_F64))) return rc;
  MH_CHECK(hipMemcpyAsync(stage, hl, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if ((rc = links_from_milc_dev(stage, L.lng, P_F64))) return rc;
  MH_CHECK(hipStreamSynchronize(c.stream));
  dev_free(stage);
  L.prec = P_F64;
  L.valid = true;
  c.links_version++;
  return 0;
}

}  // namespace

namespace mh {
// 16-bit long links in the two-row form when they are scaled-unitary (on by default)
static int g_long_recon = 1;
// relative deviation |L^dag L - c^2 1| / c^2 accepted for the two-row form: below
// the 16-bit quantisation step (2^-15), so the format adds no error of its own order
static const double kLongReconTol = 1e-5;

// Sloppy copies of the FP64 links, built on demand.
int ensure_links(int prec) {
  Context& c = ctx();
  if (c.links[prec].valid) return 0;
  if (!c.links[P_F64].valid) return -3;
  DevLinks& L = c.links[prec];
  int rc;
  if ((rc = alloc_links(L.fat, prec, c.stride)) || (rc = alloc_links(L.lng, prec, c.stride))) return rc;
  if ((rc = convert_links(c.links[P_F64].fat, P_F64, L.fat, prec))) return rc;
  L.lng.recon = 0;
  rc = 1;
  if (prec == P_H16 && g_long_recon && (rc = convert_long_h16r(c.links[P_F64].lng, L.lng, kLongReconTol)) < 0)
    return rc;
  if (rc == 1 && (rc = convert_links(c.links[P_F64].lng, P_F64, L.lng, prec))) return rc;
  L.prec = prec;
  L.valid = true;
  return 0;
}
}  // namespace mh

namespace {

// Links refresh rule of the reference glue: re-upload when the caller signals
// new links (*num_iters == -1) or passes a different link array.
int refresh_links(int ext_prec, const void* fat, const void* lng, int* num_iters) {
  Context& c = ctx();
  bool fresh = (num_iters && *num_iters == -1) || !c.links[P_F64].valid || c.link_key != fat ||
               c.link_key2 != lng;
  if (!fresh) return 0;
  int rc = upload_links(ext_prec, fat, lng);
  if (rc) return rc;
  c.link_key = fat;
  c.link_key2 = lng;
  return 0;
}

// host MILC vector (float/double) -> device FP64 parity field
int host_vec_in(int ext_prec, const void* h, VecField& v, int pb) {
  Context& c = ctx();
  if (ext_prec == 2) return upload_vec_milc((const double*)h, v, P_F64, pb);
  std::vector<double> tmp((size_t)2 * c.Vh * 6, 0.0);
  const float* f = (const float*)h;
  size_t off = (size_t)pb * c.Vh * 6;
  for (size_t i = 0; i < (size_t)c.Vh * 6; i++) tmp[off + i] = f[off + i];
  return upload_vec_milc(tmp.data(), v, P_F64, pb);
}
int host_vec_out(int ext_prec, const VecField& v, int pb, void* h) {
  Context& c = ctx();
  if (ext_prec == 2) return download_vec_milc(v, P_F64, pb, (double*)h);
  std::vector<double> tmp((size_t)2 * c.Vh * 6, 0.0);
  int rc = download_vec_milc(v, P_F64, pb, tmp.data());
  if (rc) return rc;
  float* f = (float*)h;
  size_t off = (size_t)pb * c.Vh * 6;
  for (size_t i = 0; i < (size_t)c.Vh * 6; i++) f[off + i] = (float)tmp[off + i];
  return 0;
}

int sloppy_of(int internal_precision, int mixed) {
  if (mixed == 2) return P_H16;
  if (mixed == 1) return P_F32;
  return internal_precision == 1 ? P_F32 : P_F64;
}

}  // namespace

extern "C" {

const char* milc_hip_version(void) { return "milc_hip 0.1 (gfx950)"; }

int milc_hip_init(const int dims[4], int device) {
  Context& c = ctx();
  if (c.initialised) milc_hip_finalize();
  if (dims[0] % 2 != 0 || ((long)dims[0] * dims[1] * dims[2] * dims[3]) % 2 != 0) return -2;
  for (int i = 0; i < 4; i++)
    if (dims[i] < 4) return -2;  // the Naik hop must not wrap onto itself
  MH_CHECK(hipSetDevice(device));
  c.device = device;
  int rc = scratch_init(device);
  if (rc) return rc;
  for (int i = 0; i < 4; i++) c.dims[i] = dims[i];
  c.geom = make_geom(dims);
  c.Vh = c.geom.Vh;
  c.stride = plane_stride(c.Vh);
  MH_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  c.ws.max_blocks = (c.Vh + 255) / 256 + 1024;
  rc = dev_alloc((void**)&c.ws.partials, (size_t)c.ws.max_blocks * 4 * 2 * sizeof(double));
  if (rc) return rc;
  if ((rc = dev_alloc((void**)&c.ws.scal, sizeof(CGScalars)))) return rc;
  MH_CHECK(hipHostMalloc((void**)&c.ws.scal_host, sizeof(CGScalars), hipHostMallocDefault));
  // pinned read-back space for one partial record per block (vec_norm2)
  if (!pinned_scratch((size_t)c.ws.max_blocks * 4 * sizeof(double))) return -5;
  c.initialised = true;
  return 0;
}

int milc_hip_finalize(void) {
  Context& c = ctx();
  if (!c.initialised) return 0;
  (void)hipStreamSynchronize(c.stream);
  for (int p = 0; p < 3; p++) {
    if (c.links[p].valid) {
      free_links(c.links[p].fat);
      free_links(c.links[p].lng);
      c.links[p].valid = false;
    }
  }
  for (int s = 0; s < kSlots; s++) free_vec(g_slots[s]);
  fn_table_release();
  gauge_table_release();
  batch_release();
  prn_release();
  if (c.thin_valid) free_links(c.thin);
  c.thin_valid = false;
  dev_free(c.ws.partials);
  dev_free(c.ws.scal);
  if (c.ws.scal_host) (void)hipHostFree(c.ws.scal_host);
  if (c.ws.pinned) (void)hipHostFree(c.ws.pinned);
  dev_free(c.staging);
  c.staging = nullptr;
  c.staging_elems = 0;
  c.ws = Workspace();
  dev_pool_release();
  (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.link_key = c.link_key2 = nullptr;
  c.initialised = false;
  return 0;
}

int milc_hip_invert(int external_precision, int internal_precision, double mass, milc_hip_inv_args args,
                    double target_resid, double target_relresid, const void* fatlink, const void* longlink,
                    const void* source, void* solution, double* final_resid, double* final_rel_resid,
                    int* num_iters) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  int parity = parity_from_evenodd(args.evenodd);
  if (parity < 0 || target_relresid < 0.0) return -2;
  int rc = refresh_links(external_precision, fatlink, longlink, num_iters);
  if (rc) return rc;
  int ps = sloppy_of(internal_precision, args.mixed_precision);
  if (ps != P_F64 && (rc = ensure_links(ps))) return rc;
  int pb = pbit(parity);
  VecField b, x;
  VecScope owned{&b, &x};
  if ((rc = alloc_vec(b, P_F64, c.stride)) || (rc = alloc_vec(x, P_F64, c.stride))) return rc;
  if ((rc = host_vec_in(external_precision, source, b, pb)) || (rc = host_vec_in(external_precision, solution, x, pb)))
    return rc;
  SolveParams sp;
  sp.parity = parity;
  int per = args.restart_iters > 0 ? args.restart_iters : args.max_iter;
  sp.max_iters = per > 0 ? per : 1;
  sp.max_restarts = std::max(1, (args.max_iter + sp.max_iters - 1) / sp.max_iters);
  sp.resid = target_resid;
  sp.relresid = target_relresid;
  sp.sloppy_prec = ps;
  sp.delta = kReliableDelta;
  sp.batch = 16;
  SolveStats st;
  rc = cg_solve(b, x, mass, sp, st);
  if (!rc) rc = host_vec_out(external_precision, x, pb, solution);
  if (rc) return rc;
  if (final_resid) *final_resid = std::sqrt(st.final_rsq);
  // the glue squares it into qic->final_relrsq (d_congrad5_fn_gpu.c:151), which
  // then holds my_relative_residue as on the CPU path (d_congrad5_fn_milc.c:205)
  if (final_rel_resid) *final_rel_resid = std::sqrt(st.final_relrsq);
  if (num_iters) *num_iters = st.iters;
  return st.iters;
}

int milc_hip_dslash(int external_precision, int internal_precision, milc_hip_inv_args args, const void* fatlink,
                    const void* longlink, const void* src, void* dst, int* num_iters) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  int parity = parity_from_evenodd(args.evenodd);
  if (parity < 0) return -2;
  int rc = refresh_links(external_precision, fatlink, longlink, num_iters);
  if (rc) return rc;
  int prec = internal_precision == 1 ? P_F32 : P_F64;
  if ((rc = ensure_links(prec))) return rc;
  int pb = pbit(parity);
  VecField s, d, s64, d64;
  VecScope owned{&s, &d, &s64, &d64};
  if ((rc = alloc_vec(s64, P_F64, c.stride)) || (rc = alloc_vec(d64, P_F64, c.stride))) return rc;
  if ((rc = host_vec_in(external_precision, src, s64, pb ^ 1))) return rc;
  if (prec == P_F64) {
    rc = dslash(c.links[P_F64], s64, d64, pb, P_F64);
  } else {
    if ((rc = alloc_vec(s, prec, c.stride)) || (rc = alloc_vec(d, prec, c.stride))) return rc;
    if (!(rc = convert_vec(s64, P_F64, s, prec)) && !(rc = dslash(c.links[prec], s, d, pb, prec)))
      rc = convert_vec(d, prec, d64, P_F64);
  }
  if (!rc) rc = host_vec_out(external_precision, d64, pb, dst);
  return rc;
}

int milc_hip_dev_source(uint64_t seed, int parity, int slot) {
  Context& c = ctx();
  if (!c.initialised || slot < 0 || slot >= kSlots) return -1;
  int rc;
  if (!g_slots[slot].data && (rc = alloc_vec(g_slots[slot], P_F64, c.stride))) return rc;
  return gaussian_vec_dev(seed, pbit(parity), g_slots[slot]);
}

int milc_hip_dev_read(int slot, int parity, double* host) {
  Context& c = ctx();
  if (!c.initialised || slot < 0 || slot >= kSlots || !g_slots[slot].data) return -1;
  return download_vec_milc(g_slots[slot], P_F64, pbit(parity), host);
}

int milc_hip_dev_write(int slot, int parity, const double* host) {
  Context& c = ctx();
  if (!c.initialised || slot < 0 || slot >= kSlots) return -1;
  int rc;
  if (!g_slots[slot].data && (rc = alloc_vec(g_slots[slot], P_F64, c.stride))) return rc;
  rc = upload_vec_milc(host, g_slots[slot], P_F64, pbit(parity));
  if (!rc) MH_CHECK(hipStreamSynchronize(c.stream));
  return rc;
}

// Solve on device slots; with `noalloc` every work vector comes from the
// static scratch array and nothing is allocated (the solution slot and the
// sloppy links must already exist).
static int dev_invert_impl(int slot_b, int slot_x, double mass, int parity, double resid, int mixed_precision,
                           int max_iter, double* stats, bool noalloc) {
  Context& c = ctx();
  if (!c.initialised || slot_b < 0 || slot_b >= kSlots || slot_x < 0 || slot_x >= kSlots || !g_slots[slot_b].data)
    return -1;
  int rc;
  int ps = sloppy_of(2, mixed_precision);
  if (noalloc) {
    if (parity != EVEN && parity != ODD) return -2;
    if (!g_slots[slot_x].data || !c.links[P_F64].valid || !c.links[ps].valid) return -3;
    size_t need = cg_scratch_bytes(ps, c.stride);
    if (need > kScratchBytes) {
      fprintf(stderr, "milc_hip_dev_invert_noalloc: %zu bytes of work vectors needed, scratch holds %zu\n", need,
              kScratchBytes);
      return -6;
    }
  } else {
    if (!g_slots[slot_x].data && (rc = alloc_vec(g_slots[slot_x], P_F64, c.stride))) return rc;
    if (ps != P_F64 && (rc = ensure_links(ps))) return rc;
  }
  SolveParams sp;
  sp.parity = parity;
  sp.max_iters = max_iter;
  sp.max_restarts = 5;
  sp.resid = resid;
  sp.relresid = 0;
  sp.sloppy_prec = ps;
  sp.delta = kReliableDelta;
  sp.batch = 16;
  SolveStats st;
  if (noalloc) {
    ScratchArena arena;
    if (!arena.ok) return -5;
    if ((rc = vec_zero(g_slots[slot_x], P_F64))) return rc;
    rc = cg_solve(g_slots[slot_b], g_slots[slot_x], mass, sp, st);
  } else {
    if ((rc = vec_zero(g_slots[slot_x], P_F64))) return rc;
    rc = cg_solve(g_slots[slot_b], g_slots[slot_x], mass, sp, st);
  }
  if (rc) return rc;
  if (stats) {
    stats[0] = std::sqrt(st.final_rsq);
    stats[1] = st.seconds;
    stats[2] = st.restarts;
    stats[3] = st.converged;
    stats[4] = st.iters;
    stats[5] = st.final_rsq;
  }
  return st.iters;
}

int milc_hip_dev_invert(int slot_b, int slot_x, double mass, int parity, double resid, int mixed_precision,
                        int max_iter, double* stats) {
  return dev_invert_impl(slot_b, slot_x, mass, parity, resid, mixed_precision, max_iter, stats, false);
}

int milc_hip_dev_invert_noalloc(int slot_b, int slot_x, double mass, int parity, double resid, int mixed_precision,
                                int max_iter, double* stats) {
  return dev_invert_impl(slot_b, slot_x, mass, parity, resid, mixed_precision, max_iter, stats, true);
}

int milc_hip_dev_read_links(double* fat, double* lng) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  double* stage = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  if (!(rc = links_to_milc_dev(c.links[P_F64].fat, P_F64, stage)))
    MH_CHECK(hipMemcpyAsync(fat, stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  if (!rc && !(rc = links_to_milc_dev(c.links[P_F64].lng, P_F64, stage)))
    MH_CHECK(hipMemcpyAsync(lng, stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  dev_free(stage);
  return rc;
}

int milc_hip_dev_dslash_bench(int prec, int parity, int reps, double* usec) {
  Context& c = ctx();
  if (!c.initialised || reps <= 0) return -1;
  int rc;
  if (prec < P_F64 || prec > P_H16) return -2;
  if ((rc = ensure_links(prec))) return rc;
  int pb = pbit(parity);
  VecField s, d, s64;
  VecScope owned{&s, &d, &s64};
  if ((rc = alloc_vec(s, prec, c.stride)) || (rc = alloc_vec(d, prec, c.stride))) return rc;
  if ((rc = alloc_vec(s64, P_F64, c.stride))) return rc;
  if ((rc = gaussian_vec_dev(12345, pb ^ 1, s64)) || (rc = convert_vec(s64, P_F64, s, prec))) return rc;
  for (int i = 0; i < 3; i++)
    if ((rc = dslash(c.links[prec], s, d, pb, prec))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  OnExit events([&] {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  });
  MH_CHECK(hipEventCreate(&e0));
  MH_CHECK(hipEventCreate(&e1));
  MH_CHECK(hipEventRecord(e0, c.stream));
  for (int i = 0; i < reps; i++)
    if ((rc = dslash(c.links[prec], s, d, pb, prec))) return rc;
  MH_CHECK(hipEventRecord(e1, c.stream));
  MH_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *usec = 1000.0 * ms / reps;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// link construction
// ---------------------------------------------------------------------------
extern "C" int milc_hip_dev_gauge_hisq(uint64_t seed, double scale, int* n_svd) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  for (int p = 0; p < 3; p++) {
    if (c.links[p].valid) {
      free_links(c.links[p].fat);
      free_links(c.links[p].lng);
      c.links[p].valid = false;
    }
  }
  c.link_key = c.link_key2 = nullptr;
  LinkField U;
  DevLinks& L = c.links[P_F64];
  int rc;
  if (!(rc = alloc_links(U, P_F64, c.stride)) && !(rc = alloc_links(L.fat, P_F64, c.stride)) &&
      !(rc = alloc_links(L.lng, P_F64, c.stride)) && !(rc = gauge_random_dev(seed, scale, U)) &&
      !(rc = rephase_dev(U)))
    rc = hisq_links_dev(U, L.fat, L.lng, nullptr, n_svd);
  if (rc) {
    free_links(U);
    free_links(L.fat);
    free_links(L.lng);
    return rc;
  }
  if (c.thin_valid) free_links(c.thin);
  c.thin = U;  // kept for the thin-link force
  c.thin_valid = true;
  L.prec = P_F64;
  L.valid = true;
  c.links_version++;
  return 0;
}

// Thin links from the host (MILC layout, FP32 or FP64): KS phases applied on
// the device unless phases_in, HISQ fat + Naik links built on the device; both
// become the current device links (the dev_* solvers and the thin-link force
// use them).
extern "C" int milc_hip_dev_load_thin_links(int precision, const void* thin, int phases_in, int* n_svd) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !thin) return -2;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  double* stage = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  if (precision == 2) {
    rc = hipMemcpyAsync(stage, thin, n * sizeof(double), hipMemcpyHostToDevice, c.stream) == hipSuccess ? 0 : -1;
  } else {
    std::vector<double> conv(n);
    for (size_t i = 0; i < n; i++) conv[i] = ((const float*)thin)[i];
    rc = hipMemcpyAsync(stage, conv.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream) == hipSuccess ? 0 : -1;
    if (!rc) rc = hipStreamSynchronize(c.stream) == hipSuccess ? 0 : -1;
  }
  for (int p = 0; p < 3; p++) {
    if (c.links[p].valid) {
      free_links(c.links[p].fat);
      free_links(c.links[p].lng);
      c.links[p].valid = false;
    }
  }
  c.link_key = c.link_key2 = nullptr;
  LinkField U{};
  DevLinks& L = c.links[P_F64];
  if (!rc) rc = alloc_links(U, P_F64, c.stride);
  if (!rc) rc = links_from_milc_dev(stage, U, P_F64);
  if (!rc && !phases_in) rc = rephase_dev(U);
  if (!rc && !(rc = alloc_links(L.fat, P_F64, c.stride))) rc = alloc_links(L.lng, P_F64, c.stride);
  if (!rc) rc = hisq_links_dev(U, L.fat, L.lng, nullptr, n_svd);
  (void)hipStreamSynchronize(c.stream);
  dev_free(stage);
  if (rc) {
    free_links(U);
    free_links(L.fat);
    free_links(L.lng);
    return rc;
  }
  if (c.thin_valid) free_links(c.thin);
  c.thin = U;
  c.thin_valid = true;
  L.prec = P_F64;
  L.valid = true;
  c.links_version++;
  return 0;
}

// The current device thin links in MILC layout (FP64), with the KS phases
// (with_phases) or without them (as lattice files hold them).
extern "C" int milc_hip_dev_read_thin_links(double* thin, int with_phases) {
  Context& c = ctx();
  if (!c.initialised || !c.thin_valid) return -1;
  if (!thin) return -2;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  double* stage = nullptr;
  LinkField T{};
  int rc;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  const LinkField* src = &c.thin;
  if (!with_phases) {
    // the phases are signs: applying them again removes them
    if (!(rc = alloc_links(T, P_F64, c.stride))) {
      rc = hipMemcpyAsync(T.data, c.thin.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream) ==
                   hipSuccess ? 0 : -1;
      if (!rc) rc = rephase_dev(T);
    }
    src = &T;
  }
  if (!rc) rc = links_to_milc_dev(*src, P_F64, stage);
  if (!rc) rc = hipMemcpyAsync(thin, stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream) == hipSuccess ? 0 : -1;
  (void)hipStreamSynchronize(c.stream);
  dev_free(stage);
  free_links(T);
  return rc;
}

extern "C" int milc_hip_load_hisq_links(int precision, const void* thin, void* fatlink, void* longlink, void* W_link,
                                        int* n_svd) {
  return milc_hip_load_hisq_links_naik(precision, thin, 0.0, fatlink, longlink, W_link, n_svd);
}

extern "C" int milc_hip_load_hisq_links_naik(int precision, const void* thin, double eps_naik, void* fatlink,
                                             void* longlink, void* W_link, int* n_svd) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if (precision != 1 && precision != 2) return -2;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  std::vector<double> conv;
  const double* h = (const double*)thin;
  if (precision == 1) {
    conv.resize(n);
    for (size_t i = 0; i < n; i++) conv[i] = ((const float*)thin)[i];
    h = conv.data();
  }
  LinkField U, fat, lng, W;
  double* stage = nullptr;
  OnExit cleanup([&] {
    (void)hipStreamSynchronize(c.stream);
    dev_free(stage);
    free_links(U);
    free_links(fat);
    free_links(lng);
    free_links(W);
  });
  int rc;
  if ((rc = alloc_links(U, P_F64, c.stride)) || (rc = alloc_links(fat, P_F64, c.stride)) ||
      (rc = alloc_links(lng, P_F64, c.stride)) || (rc = alloc_links(W, P_F64, c.stride)))
    return rc;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  MH_CHECK(hipMemcpyAsync(stage, h, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if ((rc = links_from_milc_dev(stage, U, P_F64))) return rc;
  if ((rc = hisq_links_dev(U, fat, lng, &W, n_svd))) return rc;
  if (eps_naik != 0.0) {
    // set = X_2 + eps_naik X_3, X_3 = (1/8 W, -1/24 W W W): the Naik paths of
    // both tables have the same coefficient, so X_3's long link is lng
    // (fermion_links_hisq_load_milc.c:597-615)
    if ((rc = link_axpy_dev(fat, W, eps_naik * kHisqNaikSet[0], 1.0, 0)) ||
        (rc = link_axpy_dev(lng, lng, eps_naik * kHisqNaikSet[1] / kHisqLevel2[1], 1.0, 0)))
      return rc;
  }
  std::vector<double> out(n);
  void* dsts[3] = {fatlink, longlink, W_link};
  LinkField* srcs[3] = {&fat, &lng, &W};
  for (int k = 0; k < 3; k++) {
    if (!dsts[k]) continue;
    if ((rc = links_to_milc_dev(*srcs[k], P_F64, stage))) return rc;
    MH_CHECK(hipMemcpyAsync(out.data(), stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MH_CHECK(hipStreamSynchronize(c.stream));
    if (precision == 2) memcpy(dsts[k], out.data(), n * sizeof(double));
    else
      for (size_t i = 0; i < n; i++) ((float*)dsts[k])[i] = (float)out[i];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// multi-shift solver and force outer products
// ---------------------------------------------------------------------------
namespace mh {
int force_oprod_coeff(const std::vector<VecField>& xe, const std::vector<VecField>& xo, const std::vector<double>& c1,
                      const std::vector<double>& c3, LinkField& o1, LinkField& o3);
}

extern "C" int milc_hip_multishift_invert(int external_precision, int internal_precision, int num_offsets,
                                          const double* offsets, milc_hip_inv_args args, const double* target_resid,
                                          const double* target_relresid, const void* fatlink, const void* longlink,
                                          const void* source, void** solutions, double* final_resid,
                                          double* final_rel_resid, int* num_iters) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  int parity = parity_from_evenodd(args.evenodd);
  if (parity < 0 || num_offsets <= 0 || num_offsets > 32) return -2;
  for (int j = 0; j < num_offsets; j++)
    if (!(offsets[j] > 0)) return -2;  // ks_multicg_offset.c:139-145
  // no Fermilab relative residual in the multi-shift solve: the reference
  // glue stops on it too (ks_multicg_offset_gpu.c:55-58)
  if (target_relresid)
    for (int j = 0; j < num_offsets; j++)
      if (target_relresid[j] != 0.0) return -4;
  int rc = refresh_links(external_precision, fatlink, longlink, num_iters);
  if (rc) return rc;
  int pb = pbit(parity);
  bool mixed = args.mixed_precision != 0;
  int bulk = (internal_precision == 1 || mixed) ? P_F32 : P_F64;
  if ((rc = ensure_links(bulk))) return rc;
  VecField b;
  std::vector<VecField> xs(num_offsets);
  VecScope owned{&b};
  for (auto& v : xs) owned.add(&v);
  if ((rc = alloc_vec(b, P_F64, c.stride))) return rc;
  for (auto& v : xs)
    if ((rc = alloc_vec(v, P_F64, c.stride))) return rc;
  if ((rc = host_vec_in(external_precision, source, b, pb))) return rc;
  std::vector<double> offs(offsets, offsets + num_offsets);
  MultiParams mp;
  mp.parity = parity;
  mp.max_iters = args.max_iter > 0 ? args.max_iter : 1000;
  // the bulk FP32 solve cannot reach FP64 targets; it stops at its own floor
  // and the per-shift mixed-precision CG below finishes each system
  mp.resid = mixed ? std::max(target_resid[0], kFp32StageResid) : target_resid[0];
  mp.prec = bulk;
  mp.batch = 16;
  SolveStats st;
  if ((rc = multicg_solve(b, xs, offs, mp, st))) return rc;
  int total_iters = st.iters;
  std::vector<double> fres(num_offsets, 0.0);
  if (mixed) {
    // per-shift refinement: (offset_j - D D) x_j = b with mass = sqrt(offset_j)/2
    // (the mass <-> offset relation of ks_multicg.c:142,222), FP32 sloppy
    // operator (16-bit with mixed_precision 2), up to four shifts per batched CG
    std::vector<double> res(num_offsets);
    for (int j = 0; j < num_offsets; j++) res[j] = target_resid[j] > 0 ? target_resid[j] : target_resid[0];
    long it = 0;
    int conv = 1;
    const int sloppy = args.mixed_precision == 2 ? P_H16 : P_F32;
    if ((rc = refine_shifts(b, xs, offs.data(), res.data(), parity, mp.max_iters, sloppy, &it, fres.data(), &conv)))
      return rc;
    total_iters += (int)it;
  } else {
    for (int j = 0; j < num_offsets; j++) fres[j] = std::sqrt(st.final_rsq);
  }
  for (int j = 0; j < num_offsets; j++)
    if ((rc = host_vec_out(external_precision, xs[j], pb, solutions[j]))) return rc;
  for (int j = 0; j < num_offsets; j++) {
    if (final_resid) final_resid[j] = fres[j];
    if (final_rel_resid) final_rel_resid[j] = 0.0;
  }
  if (num_iters) *num_iters = total_iters;
  return total_iters;
}

extern "C" int milc_hip_compute_oprod(int precision, int num_terms, double** coeff, void** quark_field,
                                      void* oprod[2]) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if (num_terms <= 0 || num_terms > 32 || (precision != 1 && precision != 2)) return -2;
  int rc;
  std::vector<VecField> xe(num_terms), xo(num_terms);
  std::vector<double> c1(num_terms), c3(num_terms);
  LinkField o1, o3;
  double* stage = nullptr;
  OnExit cleanup([&] {
    (void)hipStreamSynchronize(c.stream);
    dev_free(stage);
    free_links(o1);
    free_links(o3);
    for (int t = 0; t < num_terms; t++) {
      free_vec(xe[t]);
      free_vec(xo[t]);
    }
  });
  for (int t = 0; t < num_terms; t++) {
    c1[t] = coeff[t][0];
    c3[t] = coeff[t][1];
    if ((rc = alloc_vec(xe[t], P_F64, c.stride)) || (rc = alloc_vec(xo[t], P_F64, c.stride))) return rc;
    if ((rc = host_vec_in(precision, quark_field[t], xe[t], 0)) || (rc = host_vec_in(precision, quark_field[t], xo[t], 1)))
      return rc;
  }
  if ((rc = alloc_links(o1, P_F64, c.stride)) || (rc = alloc_links(o3, P_F64, c.stride))) return rc;
  if ((rc = force_oprod_coeff(xe, xo, c1, c3, o1, o3))) return rc;
  // device layout -> MILC AoS (4*i+dir) -> the caller's per-direction fields
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  if ((rc = dev_alloc((void**)&stage, n * sizeof(double)))) return rc;
  std::vector<double> h(n);
  LinkField* src[2] = {&o1, &o3};
  size_t V = (size_t)2 * c.Vh;
  for (int k = 0; k < 2; k++) {
    if ((rc = links_to_milc_dev(*src[k], P_F64, stage))) return rc;
    MH_CHECK(hipMemcpyAsync(h.data(), stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MH_CHECK(hipStreamSynchronize(c.stream));
    void** dirs = (void**)oprod[k];
    for (int mu = 0; mu < 4; mu++) {
      if (precision == 2) {
        double* d = (double*)dirs[mu];
        for (size_t i = 0; i < V; i++) memcpy(d + 18 * i, h.data() + 18 * (4 * i + mu), 18 * sizeof(double));
      } else {
        float* d = (float*)dirs[mu];
        for (size_t i = 0; i < V; i++)
          for (int e = 0; e < 18; e++) d[18 * i + e] = (float)h[18 * (4 * i + mu) + e];
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// measurement helpers for the benchmark
// ---------------------------------------------------------------------------
namespace {
// one 16-byte element per thread over the whole buffer: the fastest plain copy
// shape on gfx950 (the grid-stride form reaches ~75% of it)
__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_diff_norm(mh::VecField a, mh::VecField b, int Vh, double* out) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s[2] = {0, 0};
  if (cb < Vh) {
    mh::Vec3<double> x, y;
    mh::load_vec<mh::P_F64>(a, cb, x);
    mh::load_vec<mh::P_F64>(b, cb, y);
    for (int k = 0; k < 3; k++) {
      double dr = x.c[k].re - y.c[k].re, di = x.c[k].im - y.c[k].im;
      s[0] += dr * dr + di * di;
      s[1] += y.c[k].re * y.c[k].re + y.c[k].im * y.c[k].im;
    }
  }
  for (int k = 0; k < 2; k++)
    for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
  __shared__ double sm[2][16];
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = s[0];
    sm[1][threadIdx.x >> 6] = s[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      t0 += sm[0][w];
      t1 += sm[1][w];
    }
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}
}  // namespace

extern "C" int milc_hip_dev_copy_bench(size_t bytes, int reps, double* usec) {
  Context& c = ctx();
  if (!c.initialised || reps <= 0 || bytes < 16) return -1;
  size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&a, n * 16)) || (rc = dev_alloc((void**)&b, n * 16))) return rc;
  MH_CHECK(hipMemsetAsync(a, 0, n * 16, c.stream));
  int nb = (int)((n + 255) / 256);
  for (int i = 0; i < 2; i++) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c.stream, a, b, n);
  hipEvent_t e0, e1;
  MH_CHECK(hipEventCreate(&e0));
  MH_CHECK(hipEventCreate(&e1));
  MH_CHECK(hipEventRecord(e0, c.stream));
  for (int i = 0; i < reps; i++) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c.stream, a, b, n);
  MH_CHECK(hipEventRecord(e1, c.stream));
  MH_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *usec = 1000.0 * ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  return 0;
}

extern "C" int milc_hip_dev_diff(int slot_a, int slot_b, double* rel) {
  Context& c = ctx();
  if (!c.initialised || slot_a < 0 || slot_a >= kSlots || slot_b < 0 || slot_b >= kSlots || !g_slots[slot_a].data ||
      !g_slots[slot_b].data)
    return -1;
  int nb = (c.Vh + 255) / 256;
  if (nb * 2 > c.ws.max_blocks * 8) return -2;
  hipLaunchKernelGGL(k_diff_norm, dim3(nb), dim3(256), 0, c.stream, g_slots[slot_a], g_slots[slot_b], c.Vh,
                     c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h((size_t)nb * 2);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double d = 0, n = 0;
  for (int i = 0; i < nb; i++) {
    d += h[2 * i];
    n += h[2 * i + 1];
  }
  *rel = n > 0 ? std::sqrt(d / n) : 0.0;
  return 0;
}

// ---------------------------------------------------------------------------
// RHMC compositions on device slots
// ---------------------------------------------------------------------------
namespace mh {
int rhmc_action(const VecField& phi, int order, const double* roots, const double* residues, double resid, int mixed,
                double* action, int* iters);
int rhmc_force_level0(const VecField& phi, int n, const double* offsets, const double* residues, double resid,
                      int mixed, double* S, LinkField& o1, LinkField& o3, int* iters);
int force_level0_from_solutions(const std::vector<VecField>& x, const double* residues, LinkField& o1, LinkField& o3);
}

extern "C" int milc_hip_dev_rhmc_action(int slot_phi, int order, const double* roots, const double* residues,
                                        double resid, int mixed_precision, double* action) {
  Context& c = ctx();
  if (!c.initialised || slot_phi < 0 || slot_phi >= kSlots || !g_slots[slot_phi].data) return -1;
  if (order <= 0 || order > 32) return -2;
  for (int k = 1; k <= order; k++)
    if (!(roots[k] > 0)) return -2;
  int rc;
  if (!c.links[P_F64].valid) return -3;
  if (mixed_precision && (rc = ensure_links(P_F32))) return rc;
  int iters = 0;
  rc = rhmc_action(g_slots[slot_phi], order, roots, residues, resid, mixed_precision ? 1 : 0, action, &iters);
  return rc ? rc : iters;
}

extern "C" int milc_hip_dev_force_level0(int slot_phi, int n, const double* offsets, const double* residues,
                                         double resid, int mixed_precision, double* S, double* oprod1,
                                         double* oprod3) {
  Context& c = ctx();
  if (!c.initialised || slot_phi < 0 || slot_phi >= kSlots || !g_slots[slot_phi].data) return -1;
  if (n <= 0 || n > 32) return -2;
  for (int k = 0; k < n; k++)
    if (!(offsets[k] > 0)) return -2;
  int rc;
  if (!c.links[P_F64].valid) return -3;
  if (mixed_precision && (rc = ensure_links(P_F32))) return rc;
  LinkField o1, o3;
  double* stage = nullptr;
  OnExit cleanup([&] {
    (void)hipStreamSynchronize(c.stream);
    dev_free(stage);
    free_links(o1);
    free_links(o3);
  });
  if ((rc = alloc_links(o1, P_F64, c.stride)) || (rc = alloc_links(o3, P_F64, c.stride))) return rc;
  int iters = 0;
  if ((rc = rhmc_force_level0(g_slots[slot_phi], n, offsets, residues, resid, mixed_precision ? 1 : 0, S, o1, o3,
                              &iters)))
    return rc;
  if (oprod1 || oprod3) {
    size_t nn = (size_t)2 * c.Vh * 4 * 18;
    if ((rc = dev_alloc((void**)&stage, nn * sizeof(double)))) return rc;
    if (oprod1) {
      if ((rc = links_to_milc_dev(o1, P_F64, stage))) return rc;
      MH_CHECK(hipMemcpyAsync(oprod1, stage, nn * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      MH_CHECK(hipStreamSynchronize(c.stream));
    }
    if (oprod3) {
      if ((rc = links_to_milc_dev(o3, P_F64, stage))) return rc;
      MH_CHECK(hipMemcpyAsync(oprod3, stage, nn * sizeof(double), hipMemcpyDeviceToHost, c.stream));
      MH_CHECK(hipStreamSynchronize(c.stream));
    }
  }
  return iters;
}

extern "C" int milc_hip_dev_load_links(const double* fat, const double* lng) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  int rc = upload_links(2, fat, lng);
  if (rc) return rc;
  c.link_key = c.link_key2 = nullptr;  // host-array calls re-upload their own links
  return 0;
}

// ---------------------------------------------------------------------------
// FN (fat + Naik) fermion force on the thin links
// ---------------------------------------------------------------------------
namespace mh {
int fn_force_from_oprod(const LinkField& U, const LinkField& o1, const LinkField& o3, const double coeff[6],
                        double scale, double* mom_dev);
}

namespace {
// MILC AoS link array [i][dir][18] (float or double) -> FP64 device links
int host_links_in(int ext_prec, const void* h, LinkField& out, double* stage) {
  Context& c = ctx();
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  std::vector<double> conv;
  const double* src = (const double*)h;
  if (ext_prec == 1) {
    conv.resize(n);
    for (size_t i = 0; i < n; i++) conv[i] = ((const float*)h)[i];
    src = conv.data();
  }
  MH_CHECK(hipMemcpyAsync(stage, src, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  int rc = links_from_milc_dev(stage, out, P_F64);
  MH_CHECK(hipStreamSynchronize(c.stream));
  return rc;
}

// four per-direction su3_matrix[V] arrays -> FP64 device links
int host_dirs_in(int ext_prec, const void* const dirs[4], LinkField& out, double* stage) {
  Context& c = ctx();
  size_t V = (size_t)2 * c.Vh;
  std::vector<double> h(V * 4 * 18);
  for (int mu = 0; mu < 4; mu++)
    for (size_t i = 0; i < V; i++)
      for (int e = 0; e < 18; e++)
        h[18 * (4 * i + mu) + e] =
            ext_prec == 2 ? ((const double*)dirs[mu])[18 * i + e] : ((const float*)dirs[mu])[18 * i + e];
  return host_links_in(2, h.data(), out, stage);
}

// mom[(4 i + dir) * 10 + k] += increments (device, double) in the caller's precision
int add_momentum(int ext_prec, const double* dev, void* mom) {
  Context& c = ctx();
  size_t n = (size_t)2 * c.Vh * 4 * 10;
  std::vector<double> h(n);
  MH_CHECK(hipMemcpyAsync(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  if (ext_prec == 2) {
    double* m = (double*)mom;
    for (size_t i = 0; i < n; i++) m[i] += h[i];
  } else {
    float* m = (float*)mom;
    for (size_t i = 0; i < n; i++) m[i] += (float)h[i];
  }
  return 0;
}

struct ForceBuffers {
  LinkField U{}, o1{}, o3{};
  double* stage = nullptr;
  double* mom = nullptr;
  ~ForceBuffers() {
    free_links(U);
    free_links(o1);
    free_links(o3);
    dev_free(stage);
    dev_free(mom);
  }
  int alloc() {
    Context& c = ctx();
    int rc;
    if ((rc = alloc_links(U, P_F64, c.stride)) || (rc = alloc_links(o1, P_F64, c.stride)) ||
        (rc = alloc_links(o3, P_F64, c.stride)))
      return rc;
    if ((rc = dev_alloc((void**)&stage, (size_t)2 * c.Vh * 4 * 18 * sizeof(double)))) return rc;
    return dev_alloc((void**)&mom, (size_t)2 * c.Vh * 4 * 10 * sizeof(double));
  }
};
}  // namespace

extern "C" int milc_hip_asqtad_force(int precision, const double act_path_coeff[6], const void* const one_link_oprod[4],
                                     const void* const naik_oprod[4], const void* gauge, void* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !act_path_coeff || !one_link_oprod || !naik_oprod || !gauge || !momentum)
    return -2;
  ForceBuffers b;
  int rc;
  if ((rc = b.alloc())) return rc;
  if ((rc = host_links_in(precision, gauge, b.U, b.stage)) || (rc = host_dirs_in(precision, one_link_oprod, b.o1, b.stage)) ||
      (rc = host_dirs_in(precision, naik_oprod, b.o3, b.stage)))
    return rc;
  if ((rc = fn_force_from_oprod(b.U, b.o1, b.o3, act_path_coeff, 1.0, b.mom))) return rc;
  return add_momentum(precision, b.mom, momentum);
}

extern "C" int milc_hip_fermion_force_fn_multi(int precision, double eps, const double* residues, void** multi_x,
                                               int nterms, const double act_path_coeff[6], const void* gauge,
                                               void* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || nterms <= 0 || nterms > 32 || !residues || !multi_x || !act_path_coeff ||
      !gauge || !momentum)
    return -2;
  ForceBuffers b;
  int rc;
  if ((rc = b.alloc())) return rc;
  std::vector<VecField> xe(nterms), xo(nterms);
  std::vector<double> r(residues, residues + nterms);
  auto release = [&]() {
    for (int t = 0; t < nterms; t++) {
      free_vec(xe[t]);
      free_vec(xo[t]);
    }
  };
  for (int t = 0; t < nterms; t++) {
    if ((rc = alloc_vec(xe[t], P_F64, c.stride)) || (rc = alloc_vec(xo[t], P_F64, c.stride)) ||
        (rc = host_vec_in(precision, multi_x[t], xe[t], 0)) || (rc = host_vec_in(precision, multi_x[t], xo[t], 1))) {
      release();
      return rc;
    }
  }
  if ((rc = host_links_in(precision, gauge, b.U, b.stage)) || (rc = force_oprod_coeff(xe, xo, r, r, b.o1, b.o3)) ||
      (rc = fn_force_from_oprod(b.U, b.o1, b.o3, act_path_coeff, 2.0 * eps, b.mom))) {
    release();
    return rc;
  }
  rc = add_momentum(precision, b.mom, momentum);
  release();
  return rc;
}

// ---------------------------------------------------------------------------
// many right-hand sides: batched stencil, multi-shift and CG
// ---------------------------------------------------------------------------

namespace {
int batch_width(int nsrc) { return nsrc <= 4 ? 4 : (nsrc <= 8 ? 8 : 16); }

struct BBuf {
  void* p = nullptr;
  BBuf() = default;
  BBuf(const BBuf&) = delete;
  BBuf& operator=(const BBuf&) = delete;
  ~BBuf() { dev_free(p); }
};

bool slot_ok(int s, bool need_data) {
  return s >= 0 && s < kSlots && (!need_data || g_slots[s].data != nullptr);
}

int ensure_slot(int s) {
  Context& c = ctx();
  if (!g_slots[s].data) return alloc_vec(g_slots[s], P_F64, c.stride);
  return 0;
}
}  // namespace

extern "C" int milc_hip_dev_dslash(int prec, int slot_src, int parity_out, int slot_dst) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  if (prec < P_F64 || prec > P_H16 || (parity_out != EVEN && parity_out != ODD)) return -2;
  if (!slot_ok(slot_src, true) || !slot_ok(slot_dst, false) || slot_src == slot_dst) return -2;
  int rc;
  if ((rc = ensure_links(prec)) || (rc = ensure_slot(slot_dst))) return rc;
  const int pb = pbit(parity_out);
  if (prec == P_F64) return dslash(c.links[P_F64], g_slots[slot_src], g_slots[slot_dst], pb, P_F64);
  VecField s{}, d{};
  if ((rc = alloc_vec(s, prec, c.stride)) || (rc = alloc_vec(d, prec, c.stride))) {
    free_vec(s);
    return rc;
  }
  rc = convert_vec(g_slots[slot_src], P_F64, s, prec);
  if (!rc) rc = dslash(c.links[prec], s, d, pb, prec);
  if (!rc) rc = convert_vec(d, prec, g_slots[slot_dst], P_F64);
  (void)hipStreamSynchronize(c.stream);
  free_vec(s);
  free_vec(d);
  return rc;
}

extern "C" int milc_hip_dev_cg_iteration_bench(int slot_src, double mass, int parity, int reps, double* us) {
  Context& c = ctx();
  if (!c.initialised || !us || !c.links[P_F64].valid) return -1;
  if (!slot_ok(slot_src, true) || (parity != EVEN && parity != ODD)) return -2;
  int rc;
  if ((rc = ensure_links(P_H16))) return rc;
  return cg_iteration_profile(g_slots[slot_src], mass, parity, reps, us);
}

// precision code of the batched-stencil entry points: 1 float, 2 double,
// 3 the 16-bit records and links of the batched CG's sloppy operator
static int block_prec(int prec) { return prec == 2 ? P_F64 : (prec == 3 ? P_H16 : P_F32); }

extern "C" int milc_hip_dev_block_tune(int prec, int nrhs, int* variant, double* usec, char* desc, int ndesc) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  int ip = block_prec(prec);
  if (ip != P_F64 && ensure_links(ip)) return -3;
  int v;
  double us;
  int rc = btune(ip, nrhs, &v, &us);
  if (rc) return rc;
  if (variant) *variant = v;
  if (usec) *usec = us;
  if (desc && ndesc > 0) bdescribe(ip, nrhs, v, desc, ndesc);
  return (int)bvariants(ip, nrhs).size();
}

extern "C" int milc_hip_dev_block_describe(int prec, int nrhs, int variant, char* desc, int ndesc) {
  Context& c = ctx();
  if (!c.initialised || !desc || ndesc <= 0) return -1;
  int ip = block_prec(prec);
  int n = (int)bvariants(ip, nrhs).size();
  if (variant < 0 || variant >= n) return -2;
  bdescribe(ip, nrhs, variant, desc, ndesc);
  return n;
}

extern "C" int milc_hip_dev_block_dslash_bench(int prec, int nrhs, int variant, int reps, double* usec) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  int ip = block_prec(prec);
  if (ip != P_F64 && ensure_links(ip)) return -3;
  return btime_variant(ip, nrhs, variant, reps, usec);
}

extern "C" int milc_hip_dev_block_dslash(int prec, int nrhs, int variant, int nsrc, const int* src_slots,
                                         int out_parity, const int* dst_slots) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  if (nsrc <= 0 || nsrc > nrhs || (nrhs != 4 && nrhs != 8 && nrhs != 16)) return -2;
  int ip = block_prec(prec);
  int rc;
  if (ip != P_F64 && (rc = ensure_links(ip))) return rc;
  for (int s = 0; s < nsrc; s++)
    if (!slot_ok(src_slots[s], true) || !slot_ok(dst_slots[s], false)) return -2;
  BBuf a, b;
  if ((rc = balloc(&a.p, ip, nrhs)) || (rc = balloc(&b.p, ip, nrhs))) return rc;
  for (int s = 0; s < nsrc; s++)
    if ((rc = bpack(g_slots[src_slots[s]], a.p, ip, nrhs, s))) return rc;
  BDOp op;
  op.prec = ip;
  op.nr = nrhs;
  op.kind = BK_STORE;
  op.pb = pbit(out_parity);
  op.src = a.p;
  op.dst = b.p;
  op.variant = variant;
  if ((rc = bdslash(op, nullptr))) return rc;
  for (int s = 0; s < nsrc; s++) {
    if ((rc = ensure_slot(dst_slots[s]))) return rc;
    if ((rc = bunpack(b.p, ip, nrhs, s, g_slots[dst_slots[s]]))) return rc;
  }
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

extern "C" int milc_hip_dev_multishift_block(int nsrc, const int* src_slots, int num_offsets, const double* offsets,
                                             double resid, double refine_resid, int parity, int max_iters,
                                             int out_slot0, int out_nsrc, double* stats) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  if (nsrc <= 0 || nsrc > BNR_MAX || num_offsets <= 0 || num_offsets > BMS_MAX) return -2;
  if (parity != EVEN && parity != ODD) return -2;
  for (int s = 0; s < nsrc; s++)
    if (!slot_ok(src_slots[s], true)) return -2;
  if (out_slot0 >= 0 && (out_nsrc < 0 || out_nsrc > nsrc || out_slot0 + out_nsrc * num_offsets > kSlots)) return -2;
  int rc;
  if ((rc = ensure_links(P_F32))) return rc;
  const int nr = batch_width(nsrc);
  BBuf b;
  std::vector<BBuf> x(num_offsets);
  if ((rc = balloc(&b.p, P_F64, nr))) return rc;
  std::vector<void*> xp(num_offsets);
  for (int j = 0; j < num_offsets; j++) {
    if ((rc = balloc(&x[j].p, P_F64, nr))) return rc;
    xp[j] = x[j].p;
  }
  for (int s = 0; s < nsrc; s++)
    if ((rc = bpack(g_slots[src_slots[s]], b.p, P_F64, nr, s))) return rc;
  BMSParams prm;
  prm.parity = parity;
  prm.max_iters = max_iters > 0 ? max_iters : 20000;
  prm.resid = resid;
  prm.refine_resid = refine_resid;
  prm.delta = kBatchReliableDelta;
  BMSStats st;
  std::vector<double> offs(offsets, offsets + num_offsets);
  if ((rc = bmultishift(b.p, nr, offs, prm, xp, st))) return rc;
  if (out_slot0 >= 0) {
    for (int s = 0; s < out_nsrc; s++)
      for (int j = 0; j < num_offsets; j++) {
        int slot = out_slot0 + s * num_offsets + j;
        if ((rc = ensure_slot(slot)) || (rc = bunpack(x[j].p, P_F64, nr, s, g_slots[slot]))) return rc;
      }
    MH_CHECK(hipStreamSynchronize(c.stream));
  }
  if (stats) {
    int v;
    double us = 0;
    btune(P_F32, nr, &v, &us);
    stats[0] = st.ms_iters;
    stats[1] = st.ms_seconds;
    stats[2] = st.refine_seconds;
    stats[3] = st.converged;
    stats[4] = v;
    stats[5] = us;
    for (int j = 0; j < num_offsets; j++) {
      stats[6 + j] = st.true_rel[j];
      stats[6 + num_offsets + j] = st.refine_iters[j];
    }
  }
  return st.ms_iters;
}

extern "C" int milc_hip_multishift_invert_block(int external_precision, int nsrc, int num_offsets,
                                                const double* offsets, milc_hip_inv_args args,
                                                const double* target_residual, double refine_residual,
                                                const void* fatlink, const void* longlink, void* const* source,
                                                void* const* solution, double* final_residual, int* num_iters) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((external_precision != 1 && external_precision != 2) || nsrc <= 0 || num_offsets <= 0 ||
      num_offsets > BMS_MAX || !offsets || !target_residual || !source || !solution)
    return -2;
  int parity = parity_from_evenodd(args.evenodd);
  if (parity < 0) return -2;
  int rc;
  if ((rc = refresh_links(external_precision, fatlink, longlink, num_iters))) return rc;
  if ((rc = ensure_links(P_F32))) return rc;
  const int pb = pbit(parity);
  int total = 0;
  for (int s0 = 0; s0 < nsrc; s0 += BNR_MAX) {
    int ns = std::min(BNR_MAX, nsrc - s0);
    int nr = batch_width(ns);
    BBuf b;
    std::vector<BBuf> x(num_offsets);
    std::vector<void*> xp(num_offsets);
    if ((rc = balloc(&b.p, P_F64, nr))) return rc;
    for (int j = 0; j < num_offsets; j++) {
      if ((rc = balloc(&x[j].p, P_F64, nr))) return rc;
      xp[j] = x[j].p;
    }
    VecField v;
    if ((rc = alloc_vec(v, P_F64, c.stride))) return rc;
    for (int s = 0; s < ns; s++) {
      if ((rc = host_vec_in(external_precision, source[s0 + s], v, pb)) || (rc = bpack(v, b.p, P_F64, nr, s))) {
        free_vec(v);
        return rc;
      }
    }
    BMSParams prm;
    prm.parity = parity;
    prm.max_iters = args.max_iter > 0 ? args.max_iter : 20000;
    prm.resid = target_residual[0];
    prm.refine_resid = refine_residual;
    prm.delta = kBatchReliableDelta;
    BMSStats st;
    std::vector<double> offs(offsets, offsets + num_offsets);
    if ((rc = bmultishift(b.p, nr, offs, prm, xp, st))) {
      free_vec(v);
      return rc;
    }
    for (int s = 0; s < ns; s++)
      for (int j = 0; j < num_offsets; j++) {
        if ((rc = bunpack(x[j].p, P_F64, nr, s, v)) ||
            (rc = host_vec_out(external_precision, v, pb, solution[(size_t)(s0 + s) * num_offsets + j]))) {
          free_vec(v);
          return rc;
        }
        if (final_residual) final_residual[(size_t)(s0 + s) * num_offsets + j] = st.true_rel_r[j][s];
      }
    free_vec(v);
    int it = st.ms_iters;
    for (int j = 0; j < num_offsets; j++) it += st.refine_iters[j];
    total += it;
  }
  if (num_iters) *num_iters = total;
  return total;
}

extern "C" int milc_hip_invert_block(int external_precision, int nsrc, double mass, milc_hip_inv_args args,
                                     double target_residual, const void* fatlink, const void* longlink,
                                     void* const* source, void* const* solution, double* final_residual,
                                     int* num_iters) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((external_precision != 1 && external_precision != 2) || nsrc <= 0 || !source || !solution) return -2;
  int parity = parity_from_evenodd(args.evenodd);
  if (parity < 0) return -2;
  int rc;
  if ((rc = refresh_links(external_precision, fatlink, longlink, num_iters))) return rc;
  if ((rc = ensure_links(P_F32))) return rc;
  const int pb = pbit(parity);
  int total = 0;
  for (int s0 = 0; s0 < nsrc; s0 += BNR_MAX) {
    int ns = std::min(BNR_MAX, nsrc - s0);
    int nr = batch_width(ns);
    BBuf b, x;
    if ((rc = balloc(&b.p, P_F64, nr)) || (rc = balloc(&x.p, P_F64, nr))) return rc;
    VecField v;
    if ((rc = alloc_vec(v, P_F64, c.stride))) return rc;
    for (int s = 0; s < ns; s++) {
      // t_dest is the initial guess (d_congrad5_fn_gpu.c:104-110)
      if ((rc = host_vec_in(external_precision, source[s0 + s], v, pb)) || (rc = bpack(v, b.p, P_F64, nr, s)) ||
          (rc = host_vec_in(external_precision, solution[s0 + s], v, pb)) || (rc = bpack(v, x.p, P_F64, nr, s))) {
        free_vec(v);
        return rc;
      }
    }
    double off[BNR_MAX];
    for (int r = 0; r < BNR_MAX; r++) off[r] = 4.0 * mass * mass;
    BCGParams cp;
    cp.parity = parity;
    cp.max_iters = args.max_iter > 0 ? args.max_iter : 20000;
    cp.resid = target_residual;
    cp.delta = kBatchReliableDelta;
    BCGStats cs;
    if ((rc = bcg_mixed(b.p, x.p, nr, off, cp, cs))) {
      free_vec(v);
      return rc;
    }
    for (int s = 0; s < ns; s++) {
      if ((rc = bunpack(x.p, P_F64, nr, s, v)) ||
          (rc = host_vec_out(external_precision, v, pb, solution[s0 + s]))) {
        free_vec(v);
        return rc;
      }
      if (final_residual) final_residual[s0 + s] = cs.true_rel[s];
    }
    free_vec(v);
    total += cs.iters;
  }
  if (num_iters) *num_iters = total;
  return total;
}

namespace mh {
int multishift_dev(const VecField& src, int n, const double* offsets, double resid, int mixed, int parity,
                   std::vector<VecField>& x, SolveStats& st);
}

extern "C" int milc_hip_dev_multishift(int slot_src, int parity, int num_offsets, const double* offsets, double resid,
                                       int mixed_precision, int out_slot0, double* stats) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  if (!slot_ok(slot_src, true) || num_offsets <= 0 || num_offsets > 32 || (parity != EVEN && parity != ODD)) return -2;
  if (out_slot0 >= 0 && out_slot0 + num_offsets > kSlots) return -2;
  int rc;
  if (mixed_precision && (rc = ensure_links(P_F32))) return rc;
  std::vector<VecField> x;
  SolveStats st;
  rc = multishift_dev(g_slots[slot_src], num_offsets, offsets, resid, mixed_precision ? 1 : 0, parity, x, st);
  if (!rc && out_slot0 >= 0)
    for (int j = 0; j < num_offsets && !rc; j++)
      if (!(rc = ensure_slot(out_slot0 + j))) rc = vec_copy(x[j], g_slots[out_slot0 + j], P_F64);
  MH_CHECK(hipStreamSynchronize(c.stream));
  for (auto& v : x) free_vec(v);
  if (rc) return rc;
  if (stats) {
    stats[0] = st.iters;
    stats[1] = st.seconds;
    stats[2] = st.converged;
  }
  return st.iters;
}

// ---------------------------------------------------------------------------
// one RHMC molecular-dynamics force evaluation plus the FP64 action, on device
// ---------------------------------------------------------------------------
namespace {
// |a - b|^2 and |b|^2 over the valid sites of two FP64 link fields
__global__ void k_link_diff(const double2* __restrict__ a, const double2* __restrict__ b, int stride, int Vh,
                            size_t n, double* __restrict__ out) {
  double s0 = 0, s1 = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if ((int)(i % stride) >= Vh) continue;
    double2 x = a[i], y = b[i];
    double dr = x.x - y.x, di = x.y - y.y;
    s0 += dr * dr + di * di;
    s1 += y.x * y.x + y.y * y.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  __shared__ double sm[2][16];
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = s0;
    sm[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      t0 += sm[0][w];
      t1 += sm[1][w];
    }
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}

// |a - b| / |b| over n doubles
__global__ void k_flat_diff(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                            double* __restrict__ out) {
  double s0 = 0, s1 = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double d = a[i] - b[i];
    s0 += d * d;
    s1 += b[i] * b[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
  }
  __shared__ double sm[2][16];
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = s0;
    sm[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      t0 += sm[0][w];
      t1 += sm[1][w];
    }
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}

int flat_rel_diff(const double* a, const double* b, size_t n, double* rel, double* norm2_b = nullptr) {
  Context& c = ctx();
  const int nb = 1024;
  hipLaunchKernelGGL(k_flat_diff, dim3(nb), dim3(256), 0, c.stream, a, b, n, c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h(2 * nb);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double d = 0, m = 0;
  for (int i = 0; i < nb; i++) {
    d += h[2 * i];
    m += h[2 * i + 1];
  }
  *rel = m > 0 ? std::sqrt(d / m) : 0.0;
  if (norm2_b) *norm2_b = m;
  return 0;
}

int link_rel_diff(const LinkField& a, const LinkField& b, double* rel) {
  Context& c = ctx();
  const int nb = 1024;
  size_t n = (size_t)72 * a.stride;
  hipLaunchKernelGGL(k_link_diff, dim3(nb), dim3(256), 0, c.stream, (const double2*)a.data, (const double2*)b.data,
                     a.stride, c.Vh, n, c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h(2 * nb);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double d = 0, m = 0;
  for (int i = 0; i < nb; i++) {
    d += h[2 * i];
    m += h[2 * i + 1];
  }
  *rel = m > 0 ? std::sqrt(d / m) : 0.0;
  return 0;
}
}  // namespace

extern "C" int milc_hip_dev_hisq_force(int slot0, int n, const double* residues, double eps, void* momentum,
                                       double* stats) {
  Context& c = ctx();
  if (!c.initialised || !c.links[P_F64].valid) return -1;
  if (!c.thin_valid) return -3;  // needs links built on the device
  if (n <= 0 || n > 32 || !residues || !stats || slot0 < 0 || slot0 + n > kSlots) return -2;
  for (int t = 0; t < n; t++)
    if (!slot_ok(slot0 + t, true)) return -2;
  LinkField o1{}, o3{}, V{}, W{};
  double* mom = nullptr;
  auto release = [&]() {
    free_links(o1);
    free_links(o3);
    free_links(V);
    free_links(W);
    dev_free(mom);
  };
  const size_t nmom = (size_t)2 * c.Vh * 4 * 10;
  int rc;
  if ((rc = alloc_links(o1, P_F64, c.stride)) || (rc = alloc_links(o3, P_F64, c.stride)) ||
      (rc = alloc_links(V, P_F64, c.stride)) || (rc = alloc_links(W, P_F64, c.stride)) ||
      (rc = dev_alloc((void**)&mom, nmom * sizeof(double)))) {
    release();
    return rc;
  }
  MH_CHECK(hipMemsetAsync(mom, 0, nmom * sizeof(double), c.stream));
  auto t0 = std::chrono::steady_clock::now();
  rc = hisq_aux_dev(c.thin, V, W);
  (void)hipStreamSynchronize(c.stream);
  const double t_aux = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<VecField> x(g_slots + slot0, g_slots + slot0 + n);
  t0 = std::chrono::steady_clock::now();
  if (!rc) rc = force_level0_from_solutions(x, residues, o1, o3);
  (void)hipStreamSynchronize(c.stream);
  const double t_oprod = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  t0 = std::chrono::steady_clock::now();
  if (!rc) rc = hisq_force_from_oprod(c.thin, V, W, o1, o3, kHisqLevel2, kHisqLevel1, 2.0 * eps, kForceFilter, mom);
  (void)hipStreamSynchronize(c.stream);
  const double t_thin = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double rel = 0, norm2 = 0;
  if (!rc) rc = flat_rel_diff(mom, mom, nmom, &rel, &norm2);
  if (!rc && momentum) rc = add_momentum(2, mom, momentum);
  release();
  if (rc) return rc;
  stats[0] = t_oprod;
  stats[1] = t_thin;
  stats[2] = norm2;
  stats[3] = t_aux;
  return 0;
}

extern "C" int milc_hip_dev_hmc_step(int slot_phi, int n, const double* offsets, const double* residues,
                                     double resid, int mixed_precision, int compare_fp64, int thin_force,
                                     double* stats) {
  Context& c = ctx();
  if (!c.initialised || !slot_ok(slot_phi, true) || !c.links[P_F64].valid) return -1;
  if (n <= 0 || n > 32 || !stats) return -2;
  if (thin_force && !c.thin_valid) return -3;  // needs links built on the device
  for (int k = 0; k < n; k++)
    if (!(offsets[k] > 0)) return -2;
  int rc;
  if (mixed_precision && (rc = ensure_links(P_F32))) return rc;
  LinkField o1{}, o3{}, r1{}, r3{}, V{}, W{};
  double *mom = nullptr, *mom_ref = nullptr;
  auto release = [&]() {
    free_links(o1);
    free_links(o3);
    free_links(r1);
    free_links(r3);
    free_links(V);
    free_links(W);
    dev_free(mom);
    dev_free(mom_ref);
  };
  const size_t nmom = (size_t)2 * c.Vh * 4 * 10;
  if ((rc = alloc_links(o1, P_F64, c.stride)) || (rc = alloc_links(o3, P_F64, c.stride))) {
    release();
    return rc;
  }
  // force: multi-shift, D_oe, outer products (update_h_rhmc.c:77-107 with level 0)
  auto t0 = std::chrono::steady_clock::now();
  double S = 0;
  int iters = 0;
  rc = rhmc_force_level0(g_slots[slot_phi], n, offsets, residues, resid, mixed_precision ? 1 : 0, &S, o1, o3, &iters);
  (void)hipStreamSynchronize(c.stream);
  double t_force = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc) {
    release();
    return rc;
  }
  // thin-link HISQ force from the outer products (eps = 1: mom = TA(2 U F_U))
  double t_thin = 0;
  if (thin_force) {
    if ((rc = alloc_links(V, P_F64, c.stride)) || (rc = alloc_links(W, P_F64, c.stride)) ||
        (rc = dev_alloc((void**)&mom, nmom * sizeof(double))) || (rc = hisq_aux_dev(c.thin, V, W))) {
      release();
      return rc;
    }
    t0 = std::chrono::steady_clock::now();
    rc = hisq_force_from_oprod(c.thin, V, W, o1, o3, kHisqLevel2, kHisqLevel1, 2.0, kForceFilter, mom);
    (void)hipStreamSynchronize(c.stream);
    t_thin = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
      release();
      return rc;
    }
  }
  // action sum_even |a0 phi + sum_k a_k x_k|^2 (d_action_rhmc.c), a0 = 1, summed
  // in FP64 from solutions at the FP64 true-residual target (the solve in the
  // precision of the force solve)
  std::vector<double> roots(n + 1), res(n + 1);
  roots[0] = 0;
  res[0] = 1.0;
  for (int k = 0; k < n; k++) {
    roots[k + 1] = offsets[k];
    res[k + 1] = residues[k];
  }
  t0 = std::chrono::steady_clock::now();
  double action = 0;
  int it_act = 0;
  rc = rhmc_action(g_slots[slot_phi], n, roots.data(), res.data(), std::min(resid, 1e-10), mixed_precision ? 1 : 0,
                   &action, &it_act);
  (void)hipStreamSynchronize(c.stream);
  double t_action = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc) {
    release();
    return rc;
  }
  double e1 = -1, e3 = -1, em = -1, t_ref = 0, action64 = -1, ea = -1;
  if (compare_fp64) {
    int it64;
    if ((rc = rhmc_action(g_slots[slot_phi], n, roots.data(), res.data(), std::min(resid, 1e-10), 0, &action64,
                          &it64))) {
      release();
      return rc;
    }
    ea = std::fabs(action - action64) / std::fabs(action64);
    if ((rc = alloc_links(r1, P_F64, c.stride)) || (rc = alloc_links(r3, P_F64, c.stride))) {
      release();
      return rc;
    }
    double S2;
    int it2;
    t0 = std::chrono::steady_clock::now();
    rc = rhmc_force_level0(g_slots[slot_phi], n, offsets, residues, std::min(resid, 1e-12), 0, &S2, r1, r3, &it2);
    (void)hipStreamSynchronize(c.stream);
    t_ref = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!rc) rc = link_rel_diff(o1, r1, &e1);
    if (!rc) rc = link_rel_diff(o3, r3, &e3);
    if (!rc && thin_force) {
      if (!(rc = dev_alloc((void**)&mom_ref, nmom * sizeof(double))))
        rc = hisq_force_from_oprod(c.thin, V, W, r1, r3, kHisqLevel2, kHisqLevel1, 2.0, kForceFilter, mom_ref);
      if (!rc) rc = flat_rel_diff(mom, mom_ref, nmom, &em);
    }
    if (rc) {
      release();
      return rc;
    }
  }
  release();
  stats[0] = t_force;
  stats[1] = iters;
  stats[2] = S;
  stats[3] = action;
  stats[4] = t_action;
  stats[5] = it_act;
  stats[6] = e1;
  stats[7] = e3;
  stats[8] = t_ref;
  stats[9] = t_thin;
  stats[10] = em;
  stats[11] = action64;
  stats[12] = ea;
  return 0;
}

// ---------------------------------------------------------------------------
// HISQ fermion force on the thin links
// ---------------------------------------------------------------------------
namespace {
struct HisqBuffers : ForceBuffers {
  LinkField V{}, W{};
  ~HisqBuffers() {
    free_links(V);
    free_links(W);
  }
  int alloc_vw() {
    Context& c = ctx();
    int rc;
    if ((rc = alloc_links(V, P_F64, c.stride)) || (rc = alloc_links(W, P_F64, c.stride))) return rc;
    return 0;
  }
};

int quark_fields_in(int precision, void* const* multi_x, int nterms, std::vector<VecField>& xe,
                    std::vector<VecField>& xo) {
  Context& c = ctx();
  int rc;
  xe.assign(nterms, VecField{});
  xo.assign(nterms, VecField{});
  for (int t = 0; t < nterms; t++)
    if ((rc = alloc_vec(xe[t], P_F64, c.stride)) || (rc = alloc_vec(xo[t], P_F64, c.stride)) ||
        (rc = host_vec_in(precision, multi_x[t], xe[t], 0)) || (rc = host_vec_in(precision, multi_x[t], xo[t], 1)))
      return rc;
  return 0;
}
void release_fields(std::vector<VecField>& v) {
  for (auto& f : v) free_vec(f);
}
}  // namespace

extern "C" int milc_hip_hisq_force(int precision, int num_terms, int num_naik_terms, double eps, double** coeff,
                                   void** quark_field, const double level2_coeff[6], const double fat7_coeff[6],
                                   const void* w_link, const void* v_link, const void* u_link, void* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || num_terms <= 0 || num_terms > 32 || !coeff || !quark_field ||
      !level2_coeff || !fat7_coeff || !w_link || !v_link || !u_link || !momentum)
    return -2;
  if (num_naik_terms < 0 || num_naik_terms > num_terms) return -2;
  HisqBuffers b;
  int rc;
  if ((rc = b.alloc()) || (rc = b.alloc_vw())) return rc;
  if ((rc = host_links_in(precision, u_link, b.U, b.stage)) || (rc = host_links_in(precision, v_link, b.V, b.stage)) ||
      (rc = host_links_in(precision, w_link, b.W, b.stage)))
    return rc;
  std::vector<VecField> xe, xo;
  std::vector<double> c1(num_terms), c3(num_terms);
  for (int t = 0; t < num_terms; t++) {
    c1[t] = coeff[t][0];
    c3[t] = coeff[t][1];
  }
  LinkField o1n{}, o3n{};
  auto done = [&](int r) {
    release_fields(xe);
    release_fields(xo);
    free_links(o1n);
    free_links(o3n);
    return r;
  };
  if ((rc = quark_fields_in(precision, quark_field, num_terms, xe, xo)) ||
      (rc = force_oprod_coeff(xe, xo, c1, c3, b.o1, b.o3)))
    return done(rc);
  if (num_naik_terms > 0) {
    // the second Naik set: the last num_naik_terms fields again, with the
    // weights coeff[num_terms + i] = eps_naik (1/8, -1/24) 2 r_t
    // (fermion_force_hisq_multi.c:2195-2222); their three-hop products join
    // the Naik path, their one-hop products go straight onto W
    std::vector<VecField> ne(xe.end() - num_naik_terms, xe.end()), no(xo.end() - num_naik_terms, xo.end());
    std::vector<double> n1(num_naik_terms), n3(num_naik_terms);
    for (int i = 0; i < num_naik_terms; i++) {
      n1[i] = coeff[num_terms + i][0];
      n3[i] = coeff[num_terms + i][1];
    }
    if ((rc = alloc_links(o1n, P_F64, c.stride)) || (rc = alloc_links(o3n, P_F64, c.stride)) ||
        (rc = force_oprod_coeff(ne, no, n1, n3, o1n, o3n)) || (rc = link_axpy_dev(b.o3, o3n, 1.0, 1.0, 0)))
      return done(rc);
  }
  // the three-hop coefficients already carry the Naik weight
  // (fermion_force_hisq_multi.c:2190-2193): unit weight on the Naik path
  double l2[6];
  for (int i = 0; i < 6; i++) l2[i] = level2_coeff[i];
  l2[1] = 1.0;
  rc = hisq_force_from_oprod(b.U, b.V, b.W, b.o1, b.o3, l2, fat7_coeff, eps, kForceFilter, b.mom,
                             num_naik_terms > 0 ? &o1n : nullptr);
  done(0);
  if (rc) return rc;
  return add_momentum(precision, b.mom, momentum);
}

extern "C" int milc_hip_eo_fermion_force_multi(int precision, double eps, const double* residues, void** multi_x,
                                               int nterms, const void* thin_links, void* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || nterms <= 0 || nterms > 32 || !residues || !multi_x || !thin_links ||
      !momentum)
    return -2;
  HisqBuffers b;
  int rc;
  if ((rc = b.alloc()) || (rc = b.alloc_vw())) return rc;
  if ((rc = host_links_in(precision, thin_links, b.U, b.stage)) || (rc = hisq_aux_dev(b.U, b.V, b.W))) return rc;
  std::vector<VecField> xe, xo;
  std::vector<double> r(residues, residues + nterms);
  if ((rc = quark_fields_in(precision, multi_x, nterms, xe, xo)) ||
      (rc = force_oprod_coeff(xe, xo, r, r, b.o1, b.o3))) {
    release_fields(xe);
    release_fields(xo);
    return rc;
  }
  rc = hisq_force_from_oprod(b.U, b.V, b.W, b.o1, b.o3, kHisqLevel2, kHisqLevel1, 2.0 * eps, kForceFilter, b.mom);
  release_fields(xe);
  release_fields(xo);
  if (rc) return rc;
  return add_momentum(precision, b.mom, momentum);
}

// ---------------------------------------------------------------------------
// gauge action, gauge force, link update, momentum action (host arrays)
// ---------------------------------------------------------------------------
namespace {
struct GaugeBuffers {
  LinkField U{};
  double* stage = nullptr;
  double* mom = nullptr;
  ~GaugeBuffers() {
    free_links(U);
    dev_free(stage);
    dev_free(mom);
  }
  int alloc(bool with_mom) {
    Context& c = ctx();
    int rc;
    if ((rc = alloc_links(U, P_F64, c.stride)) ||
        (rc = dev_alloc((void**)&stage, (size_t)2 * c.Vh * 4 * 18 * sizeof(double))))
      return rc;
    if (with_mom && (rc = dev_alloc((void**)&mom, (size_t)2 * c.Vh * 4 * 10 * sizeof(double)))) return rc;
    return 0;
  }
};

int loop_coeffs_ok(int num_loop_types, const double* loop_coeff) {
  return num_loop_types == 3 && loop_coeff != nullptr;  // 1x1, 1x2, 1x1x1
}

// host momenta (anti_hermitmat[4 V], Real precision) -> FP64 device copy
int mom_in(int ext_prec, const void* h, double* dev) {
  Context& c = ctx();
  size_t n = (size_t)2 * c.Vh * 4 * 10;
  std::vector<double> conv;
  const double* src = (const double*)h;
  if (ext_prec == 1) {
    conv.resize(n);
    for (size_t i = 0; i < n; i++) conv[i] = ((const float*)h)[i];
    src = conv.data();
  }
  MH_CHECK(hipMemcpyAsync(dev, src, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}
}  // namespace

extern "C" int milc_hip_gauge_force(int precision, int num_loop_types, const double* loop_coeff, double eb3,
                                    const void* links, int phases_in, void* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !loop_coeffs_ok(num_loop_types, loop_coeff) || !links || !momentum)
    return -2;
  GaugeBuffers b;
  int rc;
  if ((rc = b.alloc(true)) || (rc = host_links_in(precision, links, b.U, b.stage))) return rc;
  if (phases_in && (rc = rephase_dev(b.U))) return rc;  // the phases are signs: applied again they cancel
  MH_CHECK(hipMemsetAsync(b.mom, 0, (size_t)2 * c.Vh * 4 * 10 * sizeof(double), c.stream));
  if ((rc = gauge_force_dev(b.U, loop_coeff, eb3, b.mom))) return rc;
  return add_momentum(precision, b.mom, momentum);
}

extern "C" int milc_hip_gauge_action(int precision, int num_loop_types, const double* loop_coeff, const void* links,
                                     int phases_in, double* action) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !loop_coeffs_ok(num_loop_types, loop_coeff) || !links || !action)
    return -2;
  GaugeBuffers b;
  int rc;
  if ((rc = b.alloc(false)) || (rc = host_links_in(precision, links, b.U, b.stage))) return rc;
  if (phases_in && (rc = rephase_dev(b.U))) return rc;
  return gauge_action_dev(b.U, loop_coeff, action);
}

extern "C" int milc_hip_update_u(int precision, double eps, void* links, const void* momentum, int phases_in) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !links || !momentum) return -2;
  (void)phases_in;  // exp(eps P) multiplies from the left: the sign of a phased link rides along
  GaugeBuffers b;
  int rc;
  if ((rc = b.alloc(true)) || (rc = host_links_in(precision, links, b.U, b.stage)) ||
      (rc = mom_in(precision, momentum, b.mom)) || (rc = update_u_dev(b.U, b.mom, eps)) ||
      (rc = links_to_milc_dev(b.U, P_F64, b.stage)))
    return rc;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  std::vector<double> h(n);
  MH_CHECK(hipMemcpyAsync(h.data(), b.stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  if (precision == 2) memcpy(links, h.data(), n * sizeof(double));
  else
    for (size_t i = 0; i < n; i++) ((float*)links)[i] = (float)h[i];
  return 0;
}

extern "C" int milc_hip_reunitarize(int precision, double tolerance, void* links, int phases_in,
                                    double* max_deviation, int* errors) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !links || !(tolerance > 0)) return -2;
  GaugeBuffers b;
  int rc, nerr = 0;
  double md = 0;
  if ((rc = b.alloc(false)) || (rc = host_links_in(precision, links, b.U, b.stage)) ||
      (rc = reunitarize_dev(b.U, phases_in ? 1 : 0, tolerance, &nerr, &md)) ||
      (rc = links_to_milc_dev(b.U, P_F64, b.stage)))
    return rc;
  size_t n = (size_t)2 * c.Vh * 4 * 18;
  std::vector<double> h(n);
  MH_CHECK(hipMemcpyAsync(h.data(), b.stage, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  if (precision == 2) memcpy(links, h.data(), n * sizeof(double));
  else
    for (size_t i = 0; i < n; i++) ((float*)links)[i] = (float)h[i];
  if (max_deviation) *max_deviation = md;
  if (errors) *errors = nerr;
  return 0;
}

extern "C" int milc_hip_dev_seed_rng(int iseed, int precision) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if (precision != 1 && precision != 2) return -2;
  return prn_seed_dev(iseed, precision == 1);
}

extern "C" int milc_hip_dev_ranmom(double* momentum) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if (!momentum) return -2;
  size_t n = (size_t)2 * c.Vh * 4 * 10;
  double* mom = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&mom, n * sizeof(double)))) return rc;
  if (!(rc = ranmom_prn_dev(mom))) {
    rc = hipMemcpyAsync(momentum, mom, n * sizeof(double), hipMemcpyDeviceToHost, c.stream) == hipSuccess ? 0 : -1;
    (void)hipStreamSynchronize(c.stream);
  }
  dev_free(mom);
  return rc;
}

extern "C" int milc_hip_dev_gaussian_source(int parity, int slot, double* sum) {
  Context& c = ctx();
  if (!c.initialised || slot < 0 || slot >= kSlots) return -1;
  if (parity != EVEN && parity != ODD) return -2;
  int rc;
  if (!g_slots[slot].data && (rc = alloc_vec(g_slots[slot], P_F64, c.stride))) return rc;
  return gaussian_prn_dev(pbit(parity), g_slots[slot], sum);
}

extern "C" int milc_hip_mom_action(int precision, const void* momentum, double* action) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if ((precision != 1 && precision != 2) || !momentum || !action) return -2;
  double* mom = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&mom, (size_t)2 * c.Vh * 4 * 10 * sizeof(double)))) return rc;
  if (!(rc = mom_in(precision, momentum, mom))) rc = mom_action_dev(mom, action);
  dev_free(mom);
  return rc;
}

extern "C" int milc_hip_debug_fail_trajectory(int stage) {
  if (stage < 0 || stage > 3) return -2;
  rhmc_fail_at(stage);
  return 0;
}

extern "C" int milc_hip_dev_rhmc_trajectory(const milc_hip_rhmc_params* p, double* stats) {
  Context& c = ctx();
  if (!c.initialised) return -1;
  if (!p || !stats) return -2;
  if (p->n_pseudo <= 0 || p->n_pseudo > MILC_HIP_MAX_PSEUDO) return -2;
  static_assert(MILC_HIP_MAX_PSEUDO <= RHMC_MAX_PSEUDO, "pseudofermion limits");
  RhmcParams q{};
  q.steps = p->steps;
  q.integrator = p->integrator;
  q.epsilon = p->epsilon;
  q.beta = p->beta;
  for (int k = 0; k < 3; k++) q.loop_coeff[k] = p->loop_coeff[k];
  q.action = p->action;
  q.u0 = p->u0;
  if (q.action != 0 && !(q.action == 1 && q.u0 > 0)) return -2;
  q.n_pseudo = p->n_pseudo;
  for (int i = 0; i < p->n_pseudo; i++) {
    const milc_hip_pseudofermion& a = p->pf[i];
    RhmcPseudo& b = q.pf[i];
    b.mass = a.mass;
    const milc_hip_rational* src[3] = {&a.md, &a.gr, &a.fa};
    RhmcRat* dst[3] = {&b.md, &b.gr, &b.fa};
    for (int r = 0; r < 3; r++) {
      dst[r]->order = src[r]->order;
      memcpy(dst[r]->res, src[r]->res, sizeof dst[r]->res);
      memcpy(dst[r]->pole, src[r]->pole, sizeof dst[r]->pole);
    }
    b.resid_md = a.resid_md;
    b.resid_gr = a.resid_gr;
    b.resid_fa = a.resid_fa;
    b.naik_eps = a.naik_epsilon;
    if (b.naik_eps != 0.0 && q.action != 0) return -2;  // Naik sets are HISQ's
  }
  q.mixed_md = p->mixed_md;
  q.mixed_action = p->mixed_action;
  q.action_from_md = p->action_from_md;
  q.accept = p->accept;
  q.site_rng = p->site_rng;
  q.seed = p->seed;
  if (q.accept < 0 || q.accept > 2) return -2;
  return rhmc_trajectory(q, stats);
}

extern "C" int milc_hip_set_link_compression(int on) {
  Context& c = ctx();
  if (on != 0 && on != 1) return -2;
  if (on != g_long_recon && c.initialised && c.links[P_H16].valid) {
    // rebuilt in the other format on next use
    free_links(c.links[P_H16].fat);
    free_links(c.links[P_H16].lng);
    c.links[P_H16].valid = false;
  }
  g_long_recon = on;
  return 0;
}

extern "C" int milc_hip_link_compression(int prec) {
  Context& c = ctx();
  if (!c.initialised || prec < P_F64 || prec > P_H16) return -1;
  int rc;
  if ((rc = ensure_links(prec))) return rc;
  return c.links[prec].lng.recon;
}

extern "C" int milc_hip_set_force_form(int form) {
  if (form < 0 || form > 2) return -2;
  set_force_form(form);
  return 0;
}

// libspinach_powder: C-ABI, host-side planning and launches.
//
// Host planning (once per call, on the union sparsity pattern of I and all
// Q slots) replaces the per-orientation reduce(L',coil) of
// kernel/evolution.m:141-145: the coil-connected component of the union
// pattern is an invariant subspace of every L(Omega) that contains every state
// the coil can see, so restricting to it is exact (the reference's zero track
// elimination and path tracing, kernel/utilities/zte.m and path_trace.m, keep
// the same component up to their drop thresholds).
//
// Everything per orientation runs on the GPU (powder_kernels.hpp). There is
// no CPU compute path: a missing device is an error.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spinach_powder.h"
#include "powder_kernels.hpp"
#include "mas_kernels.hpp"
#include "spectral_kernels.hpp"
#include "step_kernels.hpp"
#include "planner.hpp"
#include "builder.hpp"

// Device memory of sp_powder_fid_noalloc, which calls no allocator: the staged
// inputs, per-item spectra, per-orientation FIDs, orientation-sum partials and
// work-queue heads of one call. Every device the library runs on holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the device (sp_create records it).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static_assert(spk::kSell == spp::kSellWave, "slice row width of the stepping kernel and the planner");
static_assert(spk::kStepMaxDim == spp::kStepMaxDim, "largest dimension of the stepping path");

namespace {

using namespace spp;

thread_local std::string g_create_err;

#define HIPCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      throw SpError(SP_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw SpError(SP_ENOMEM, std::string("hipMalloc of ") + std::to_string(b) + " bytes: " + hipGetErrorString(e));
    }
    bytes = b;
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

// Blocking copy: planning is not timed, and the host vectors may be temporaries.
template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) {
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
}

}  // namespace

struct sp_ctx {
  int device = 0;
  void* scratch = nullptr;  // this device's copy of the static scratch (sp_powder_fid_noalloc)
  std::string scratch_err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  hipEvent_t evs[4] = {nullptr, nullptr, nullptr, nullptr};  // stage boundaries inside the propagation
  int n_cu = 0;
  int code = 0;
  std::string err;
};

struct sp_plan {
  sp_ctx* ctx = nullptr;
  int mode = SP_MODE_FP64;
  int64_t n_systems = 0;
  int64_t nsteps = 0;
  int64_t n_items = 0;
  int64_t total_or = 0;
  double dt = 0;
  int32_t np_max = 16, np_prop = 16, block = 1, nblk = 1, log2b = 0, taylor_k = 18, ps_q = 4;
  double theta = 1.0;
  int grid = 1;
  int per_cu = 1;
  int64_t work_stride = 0;
  std::vector<int64_t> or_off, n_or;
  std::vector<double*> fid_targets;
  sp_stats last{};
  int64_t full_dim = 0, red_dim = 0, red_nnz = 0;
  int32_t real_systems = 0;
  // spectral path (real generators, FP64, padded dimension <= spk::kTriGNP); the first
  // n_spec_w spectral items have a padded dimension above spk::kTriNP (k_tridiag_w), the first
  // n_spec_g of them above spk::kTriWNP (with its global-memory prologue)
  int64_t n_prop = 0, n_spec = 0, n_spec_w = 0, n_spec_g = 0;
  int spec_nm = spk::kTriNP;  // per-item spectrum layout: 6 vectors of spec_nm doubles
  DevBuf d_tri_work;          // the prologue's workspace, spk::kTriGWork doubles per workgroup
  int64_t n_prop_complex = 0;  // the first n_prop_complex propagator items have a complex generator
  int32_t spectral_systems = 0;
  // stepping path (SP_MODE_KRYLOV, dimensions >= krylov_tol or above the dense limit)
  int64_t n_kry = 0, n_kry_complex = 0;  // the first n_kry_complex stepping items have a complex generator
  int32_t krylov_systems = 0;
  int32_t step_vec_words = 0;            // LDS doubles of the four state vectors
  int32_t step_max_sell = 0;             // largest slice layout over the plan's systems
  int32_t step_vec_global = 0;           // rho and the coil copy in the workgroup's global slot
  int64_t step_gvec_off = 0;             // their offset in the slot (doubles)
  bool step_vl[2] = {false, false};      // [REAL] values staged in LDS
  int step_rcs = 0;                      // register-cached product: slices per wave (1, 2; 0: off)
  size_t step_lds[2] = {0, 0};           // [REAL] dynamic LDS bytes
  int grid_step = 0, step_per_cu = 0;
  size_t step_work_words = 0;
  int tri_per_cu = 2, syn_per_cu = 3;  // spectral path: resident workgroups per CU
  DevBuf d_sdesc, d_sell_start, d_sell_col, d_sell_pos, d_sell_perm, d_nsub, d_terms, d_step_work;
  double plan_ms = 0.0;  // host planning and upload (reported in every sp_stats)
  // dead time (acquire.m:65-67): per-orientation initial states
  double dead_time = 0.0;
  int64_t rho_or_stride = 0;
  DevBuf d_rho_or;
  // per-orientation initial states set by the caller (sp_plan_set_initial_states), and the
  // states written by sp_plan_propagate; single-system plans keep their reduced-state map
  bool init_states = false;
  std::vector<int64_t> red_states;
  DevBuf d_prop_out, d_nsub_fs;  // propagated states; substeps of the final-state passes
  bool use_rho_or() const { return dead_time > 0 || init_states; }
  int grid_tri = 0, grid_syn = 0;
  DevBuf d_spec, d_status;
  // orientation sum: chunks of consecutive orientations, chunk ranges per system
  int64_t n_chunks = 0;
  DevBuf d_chunks, d_chunk_off, d_partial;
  DevBuf d_sys, d_items, d_queue, d_row, d_col, d_tpos, d_vals, d_slots, d_rho, d_coil;
  int32_t relax_systems = 0;  // systems with a relaxation superoperator (sp_plan_powder_relax)
  std::vector<spk::Item> h_items;  // the device item list, in queue order
  std::vector<int32_t> h_np;       // padded dimension of every system
  int dc_max_iter = 0;             // secular evaluations per root of the eigensolver's second pass (0: default)
  int32_t eig_fallbacks = 0;       // spectral orientations of the last pass re-run on the propagator path
  DevBuf d_rvals, d_rdiag;
  DevBuf d_al, d_be, d_ga, d_w, d_work, d_per, d_fid, d_sq;
  bool executed = false;
  bool graph = false;                // SP_MODE_GRAPH: the pass is recorded once and replayed
  hipGraphExec_t gexec = nullptr;
  // magic-angle spinning plans
  bool mas = false;
  int32_t ticks = 0, n_rho = 0, mas_np = 0, mas_nnz = 0, mas_nq = 0;
  DevBuf d_rot2lab, d_lab2rot;
  // tiled MAS (FP64, every invariant component of the union pattern <= 16 states): tiles, the
  // pattern's tile-local positions and per-tile ranges, the tile pairs of coil and rho0
  bool mas_tiled = false;
  int32_t mas_ntile = 0, mas_npair = 0, mas_ldet = 0;  // ldet: detected elements, padded to 16
  bool mas_coil_real = false;
  DevBuf d_epos, d_toff, d_pair_s, d_pair_t, d_coil_t, d_rho0_t, d_emap, d_nprod;
};

namespace {

int fail(sp_ctx* ctx, int code, const std::string& msg) {
  if (ctx) {
    ctx->err = msg;
    ctx->code = code;
  }
  return code;
}

// Chunk table and partial-sum buffer of the two-pass orientation sum.
void setup_sum(sp_plan* pl, hipStream_t st) {
  if (pl->n_systems > 65535) throw SpError(SP_ERANGE, "at most 65535 systems per plan");
  std::vector<spk::SumChunk> chunks;
  std::vector<int64_t> off(1, 0);
  for (int64_t s = 0; s < pl->n_systems; s++) {
    for (int64_t k = 0; k < pl->n_or[s]; k += spk::kSumChunk)
      chunks.push_back(spk::SumChunk{pl->or_off[s] + k, (int32_t)std::min<int64_t>(spk::kSumChunk, pl->n_or[s] - k), 0});
    off.push_back((int64_t)chunks.size());
  }
  if (chunks.size() > 65535) throw SpError(SP_ERANGE, "too many orientations for one plan");
  pl->n_chunks = (int64_t)chunks.size();
  upload(pl->d_chunks, chunks, st);
  upload(pl->d_chunk_off, off, st);
  pl->d_partial.alloc((size_t)std::max<int64_t>(1, pl->n_chunks) * (pl->nsteps + 1) * 2 * sizeof(double));
}

// Time-block length B = 2^l of the two-level power scheme, from a cost model
// in units of one np x np complex product: l squarings, the U-block doubling
// steps (a step with 2^k rows costs ceil(2^k/64)/ceil(np/64) of a product),
// and ceil(npts/B)-1 matrix-vector steps of the V chain. A V step streams the
// np x np matrix from memory; measured on MI355X at np = 192 it costs 0.14 of
// a product, and the ratio scales as 1/np (memory ~ np^2, product ~ np^3).
int choose_log2_block(int64_t npts, int np) {
  const double v_step = 0.14 * 192.0 / (double)np;
  const double row_tiles = std::ceil(np / 64.0);
  int best = 0;
  double best_cost = 1e300;
  for (int l = 0; l <= 10; l++) {
    const int64_t B = (int64_t)1 << l;
    if (B > npts && l > 0) break;
    double cost = l;
    for (int k = 0; k < l; k++) cost += std::ceil(std::ldexp(1.0, k) / 64.0) / row_tiles;
    cost += (double)((npts + B - 1) / B - 1) * v_step;
    if (cost < best_cost) {
      best_cost = cost;
      best = l;
    }
  }
  return best;
}

void apply_cached_tuning(sp_plan* pl);

// Propagator path: time block B = 2^log2b of the power scheme and resident
// workgroups per CU (the workspace holds one slot per workgroup).
void configure_prop(sp_ctx* ctx, sp_plan* pl, int log2b, int per_cu) {
  const int64_t npts = pl->nsteps + 1;
  pl->log2b = std::max(0, std::min(10, log2b));
  while (pl->log2b > 0 && ((int64_t)1 << pl->log2b) > npts) pl->log2b--;
  pl->block = 1 << pl->log2b;
  pl->nblk = (int32_t)((npts + pl->block - 1) / pl->block);
  pl->per_cu = std::max(1, std::min(8, per_cu));
  pl->grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * pl->per_cu, pl->n_prop));
  const size_t esz = (pl->mode == SP_MODE_FP64) ? sizeof(double) : sizeof(float);
  const int64_t np = pl->np_prop, mat = 2 * np * np;
  pl->work_stride = 6 * mat + 2 * (int64_t)(pl->block + 16) * np + 2 * (int64_t)pl->nblk * np;
  pl->work_stride = (pl->work_stride + 63) / 64 * 64;
  pl->d_work.alloc((size_t)pl->grid * pl->work_stride * esz);
}

// Spectral path: resident tridiagonalisation and synthesis workgroups per CU.
void configure_spec(sp_ctx* ctx, sp_plan* pl, int tri_per_cu, int syn_per_cu) {
  pl->tri_per_cu = std::max(1, std::min(4, tri_per_cu));
  pl->syn_per_cu = std::max(1, std::min(8, syn_per_cu));
  pl->grid_tri = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)pl->tri_per_cu * ctx->n_cu, pl->n_spec));
  pl->grid_syn = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)pl->syn_per_cu * ctx->n_cu, pl->n_spec));
}

// LDS budget of the stepping kernel per variant: the four state vectors, and the
// slice-layout values and columns when the plan's largest layout fits beside
// them; otherwise the values go to a global slot per workgroup.
constexpr size_t kStepLdsMax = 156 * 1024;

template <bool REAL, bool VL, int RCS = 0>
void step_attr(size_t bytes) {
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&spk::k_step<REAL, VL, RCS>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void configure_step(sp_ctx* ctx, sp_plan* pl, int per_cu);

void setup_step(sp_ctx* ctx, sp_plan* pl) {
  const size_t vec_bytes = (size_t)pl->step_vec_words * sizeof(double);
  if (vec_bytes > kStepLdsMax) throw SpError(SP_ERANGE, "state vectors exceed the stepping path's LDS budget");
  size_t work_words = 0;
  for (int real = 0; real < 2; real++) {
    const size_t vw = real ? 1 : 2;
    const size_t with_vals = vec_bytes + (size_t)pl->step_max_sell * (vw * sizeof(double) + sizeof(int32_t));
    pl->step_vl[real] = with_vals <= kStepLdsMax;
    pl->step_lds[real] = pl->step_vl[real] ? with_vals : vec_bytes;
    if (!pl->step_vl[real]) work_words = std::max(work_words, (vw * (size_t)pl->step_max_sell + 1) / 2 * 2);
  }
  // values in global memory for both kinds and one workgroup per CU with four vectors in LDS:
  // rho and the coil copy move to the workgroup's slot when that fits a second workgroup
  pl->step_vec_global = 0;
  pl->step_gvec_off = 0;
  const size_t cu_lds = 160 * 1024, wg_static = 1024;
  if (!pl->step_vl[0] && !pl->step_vl[1] && 2 * (vec_bytes + wg_static) > cu_lds &&
      2 * (vec_bytes / 2 + wg_static) <= cu_lds && std::getenv("SPINACH_STEP_VEC_LDS") == nullptr) {
    pl->step_vec_global = 1;
    pl->step_gvec_off = (int64_t)work_words;
    work_words += (size_t)pl->step_vec_words / 2;  // two of the four vectors
    pl->step_lds[0] = pl->step_lds[1] = vec_bytes / 2;
  }
  // both variants of each kind may run in one pass
  if (pl->step_vl[0]) step_attr<false, true>(pl->step_lds[0]); else step_attr<false, false>(pl->step_lds[0]);
  if (pl->step_vl[1]) step_attr<true, true>(pl->step_lds[1]); else step_attr<true, false>(pl->step_lds[1]);
  if (pl->step_rcs == 1) {
    if (pl->step_vl[0]) step_attr<false, true, 1>(pl->step_lds[0]);
    if (pl->step_vl[1]) step_attr<true, true, 1>(pl->step_lds[1]);
  } else if (pl->step_rcs == 2) {
    if (pl->step_vl[0]) step_attr<false, true, 2>(pl->step_lds[0]);
    if (pl->step_vl[1]) step_attr<true, true, 2>(pl->step_lds[1]);
  }
  pl->step_work_words = work_words;
  configure_step(ctx, pl, 0);
}

// Resident stepping workgroups per CU (0: as many as the LDS holds, at most 4).
void configure_step(sp_ctx* ctx, sp_plan* pl, int per_cu) {
  const size_t lds_max = std::max(pl->step_lds[0], pl->step_lds[1]);
  const int fit = std::max<int>(1, std::min<int>(4, (int)((160 * 1024) / (lds_max + 1024))));
  if (per_cu <= 0) per_cu = fit;
  pl->step_per_cu = std::min(per_cu, fit);
  pl->grid_step = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * pl->step_per_cu, pl->n_items));
  pl->d_step_work.alloc(std::max<size_t>(1, pl->step_work_words) * pl->grid_step * sizeof(double));
}

sp_plan* build_plan(sp_ctx* ctx, std::vector<SysPlan>& sps, double dt, int64_t nsteps, int mode, const Tols& tl) {
  const bool force_complex = (mode & SP_MODE_COMPLEX) != 0;
  const bool force_prop = force_complex || (mode & SP_MODE_PROPAGATOR) != 0;
  const bool force_krylov = (mode & SP_MODE_KRYLOV) != 0;
  const bool graph = (mode & SP_MODE_GRAPH) != 0;
  mode &= ~(SP_MODE_COMPLEX | SP_MODE_PROPAGATOR | SP_MODE_GRAPH | SP_MODE_KRYLOV);
  if (!(std::isfinite(dt)) || dt <= 0) throw SpError(SP_EINVAL, "dt must be positive and finite");
  if (nsteps < 0) throw SpError(SP_EINVAL, "nsteps must be non-negative");
  if (nsteps > (int64_t)1 << 30) throw SpError(SP_ERANGE, "nsteps is too large");
  if (mode != SP_MODE_FP64 && mode != SP_MODE_FP32) throw SpError(SP_EINVAL, "unknown precision mode");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<sp_plan> pl(new sp_plan);
  pl->ctx = ctx;
  pl->mode = mode;
  pl->graph = graph;
  pl->dt = dt;
  pl->nsteps = nsteps;
  pl->n_systems = (int64_t)sps.size();
  const int64_t npts = nsteps + 1;
  // path of each system: spectral (real generator, FP64, fits the register
  // layout of the tridiagonalisation) or the propagator power scheme
  // stepping: forced (sys.enable 'krylov'), at krylov_tol (evolution.m:531-533), or above the dense limit
  std::vector<char> spectral(sps.size(), 0), krylov(sps.size(), 0);
  int np_prop = 16, np_spec = 16;
  for (size_t s = 0; s < sps.size(); s++) {
    krylov[s] = (force_krylov || (double)sps[s].n >= tl.krylov_tol || sps[s].np > kDenseMaxDim) ? 1 : 0;
    spectral[s] =
        (!krylov[s] && !force_prop && mode == SP_MODE_FP64 && sps[s].real && sps[s].np <= spk::kTriGNP) ? 1 : 0;
    if (spectral[s])
      np_spec = std::max(np_spec, sps[s].np);
    else if (!krylov[s])
      np_prop = std::max(np_prop, sps[s].np);
  }
  pl->log2b = choose_log2_block(npts, np_prop);
  if (const char* e = getenv("SPINACH_LOG2_BLOCK")) pl->log2b = std::max(0, std::min(10, atoi(e)));
  pl->block = 1 << pl->log2b;
  pl->nblk = (int32_t)((npts + pl->block - 1) / pl->block);
  pl->np_prop = np_prop;
  // Taylor degree K for |X|_1 <= theta = 1: the remainder is below
  // 1.05/(K+1)!, i.e. 8.6e-18 for K=18 (FP64) and 2.6e-8 for K=10 (FP32).
  pl->theta = 1.0;
  if (mode == SP_MODE_FP64) {
    pl->taylor_k = 18;
    pl->ps_q = 4;  // X^2..X^4 + 4 Horner steps = 7 products
  } else {
    pl->taylor_k = 10;
    pl->ps_q = 3;  // X^2, X^3 + 3 Horner steps = 5 products
  }
  // the register-cached stepping product: one configuration (slices per wave) that every
  // system of the plan fits, the fewest slices per wave first
  for (int rcs = 1; rcs <= 2 && pl->step_rcs == 0; rcs++) {
    bool all = true;
    for (auto& sp : sps) all = all && sell_lanes_cached(sp, rcs, spk::kRcEntries) > 0;
    if (all) pl->step_rcs = rcs;
  }
  std::vector<spk::SysDesc> descs;
  std::vector<spk::StepDesc> sdescs;
  std::vector<int32_t> sell_start, sell_col, sell_pos, sell_perm;
  std::vector<int32_t> row, col, tpos, slots;
  std::vector<double> vals, rho, coil, al, be, ga, w, rvals, rdiag;
  // queue order: longest-processing-time first within each path (stable within a system)
  std::vector<std::pair<double, spk::Item>> prop_items, spec_items, kry_items;
  int32_t nmax = 1;
  int64_t or_base = 0;
  for (size_t s = 0; s < sps.size(); s++) {
    SysPlan& sp = sps[s];
    spk::SysDesc d{};
    d.n = sp.n;
    d.np = sp.np;
    d.nnz = sp.nnz;
    d.nq = (int32_t)sp.slots.size();
    d.pat_off = (int64_t)row.size();
    d.val_off = (int64_t)vals.size() / 2;
    d.slot_off = (int64_t)slots.size();
    d.vec_off = (int64_t)rho.size() / 2;
    d.or_off = or_base;
    d.n_or = (int64_t)sp.weights.size();
    d.real = (sp.real && !force_complex) ? 1 : 0;
    d.rlx_off = -1;
    d.rdiag_off = -1;
    if (!sp.rvals.empty()) {
      d.rlx_off = (int64_t)rvals.size() / 2;
      rvals.insert(rvals.end(), sp.rvals.begin(), sp.rvals.end());
    }
    if (!sp.rdiag.empty()) {
      d.rdiag_off = (int64_t)rdiag.size();
      rdiag.insert(rdiag.end(), sp.rdiag.begin(), sp.rdiag.end());
    }
    pl->relax_systems += sp.relax ? 1 : 0;
    for (int k = 0; k < spk::kRowBlocks; k++)
      d.rb[k] = (int32_t)(std::lower_bound(sp.row.begin(), sp.row.end(), 16 * k) - sp.row.begin());
    pl->real_systems += d.real;
    pl->spectral_systems += spectral[s];
    pl->krylov_systems += krylov[s];
    descs.push_back(d);
    pl->h_np.push_back(sp.np);
    SellDesc sl{};
    // lanes per row: the register-cached product's layout when the plan uses it
    // states in reverse Cuthill-McKee order: neighbouring rows gather neighbouring words
    const std::vector<int32_t> order = rcm_order(sp);
    const int64_t perm_off = (int64_t)sell_perm.size();
    sell_perm.insert(sell_perm.end(), order.begin(), order.end());
    sell_layout(sp, sell_start, sell_col, sell_pos, sl,
                pl->step_rcs > 0 ? sell_lanes_cached(sp, pl->step_rcs, spk::kRcEntries) : 0, order);
    int lg = 0;
    while ((1 << lg) < sl.lanes) lg++;
    if (sl.nslices > spk::kStepMaxSlices) throw SpError(SP_ERANGE, "too many slices for the stepping path");
    int32_t wmax = 0;
    for (int32_t q = 0; q < sl.nslices; q++)
      wmax = std::max(wmax, (sell_start[sl.slice_off + q + 1] - sell_start[sl.slice_off + q]) >> 6);
    spk::StepDesc sdd{sl.nslices, sl.sell_nnz, sl.slice_off, sl.sell_off, sl.lanes, lg, wmax, 0, perm_off};
    sdescs.push_back(sdd);
    pl->step_max_sell = std::max(pl->step_max_sell, sdd.sell_nnz);
    nmax = std::max(nmax, sp.n);
    row.insert(row.end(), sp.row.begin(), sp.row.end());
    col.insert(col.end(), sp.col.begin(), sp.col.end());
    tpos.insert(tpos.end(), sp.tpos.begin(), sp.tpos.end());
    slots.insert(slots.end(), sp.slots.begin(), sp.slots.end());
    vals.insert(vals.end(), sp.vals.begin(), sp.vals.end());
    rho.insert(rho.end(), sp.rho.begin(), sp.rho.end());
    coil.insert(coil.end(), sp.coil.begin(), sp.coil.end());
    al.insert(al.end(), sp.alphas.begin(), sp.alphas.end());
    be.insert(be.end(), sp.betas.begin(), sp.betas.end());
    ga.insert(ga.end(), sp.gammas.begin(), sp.gammas.end());
    w.insert(w.end(), sp.weights.begin(), sp.weights.end());
    pl->or_off.push_back(or_base);
    pl->n_or.push_back(d.n_or);
    // cost per orientation: dense stages ~ np^3, stepping ~ slice-layout entries
    const double c = krylov[s] ? (double)sdd.sell_nnz : (double)sp.np * sp.np * sp.np;
    auto& dst = krylov[s] ? kry_items : (spectral[s] ? spec_items : prop_items);
    for (int64_t k = 0; k < d.n_or; k++) dst.emplace_back(c, spk::Item{(int32_t)s, 0, or_base + k});
    or_base += d.n_or;
    pl->np_max = std::max(pl->np_max, sp.np);
    pl->full_dim = std::max<int64_t>(pl->full_dim, sp.full_dim);
    pl->red_dim = std::max<int64_t>(pl->red_dim, sp.n);
    pl->red_nnz = std::max<int64_t>(pl->red_nnz, sp.nnz);
  }
  auto by_cost = [](const std::pair<double, spk::Item>& a, const std::pair<double, spk::Item>& b) {
    return a.first > b.first;
  };
  // propagator items: complex generators first, then real ones (two launches)
  std::stable_sort(prop_items.begin(), prop_items.end(),
                   [&](const std::pair<double, spk::Item>& a, const std::pair<double, spk::Item>& b) {
                     const int ra = descs[a.second.sys].real, rb = descs[b.second.sys].real;
                     if (ra != rb) return ra < rb;
                     return a.first > b.first;
                   });
  // spectral items by size class (above kTriWNP, above kTriNP, the rest: the eigensolver and the
  // tridiagonalisation are sized per class), then by cost
  auto spec_class = [&](const spk::Item& it) {
    const int np = descs[it.sys].np;
    return np > spk::kTriWNP ? 2 : (np > spk::kTriNP ? 1 : 0);
  };
  std::stable_sort(spec_items.begin(), spec_items.end(),
                   [&](const std::pair<double, spk::Item>& a, const std::pair<double, spk::Item>& b) {
                     const int wa = spec_class(a.second), wb = spec_class(b.second);
                     if (wa != wb) return wa > wb;
                     return a.first > b.first;
                   });
  // stepping items: complex generators first, then real ones (two launches)
  std::stable_sort(kry_items.begin(), kry_items.end(),
                   [&](const std::pair<double, spk::Item>& a, const std::pair<double, spk::Item>& b) {
                     const int ra = descs[a.second.sys].real, rb = descs[b.second.sys].real;
                     if (ra != rb) return ra < rb;
                     return a.first > b.first;
                   });
  pl->n_prop_complex = 0;
  for (auto& c : prop_items)
    if (!descs[c.second.sys].real) pl->n_prop_complex++;
  pl->n_kry_complex = 0;
  for (auto& c : kry_items)
    if (!descs[c.second.sys].real) pl->n_kry_complex++;
  std::vector<spk::Item> ordered;
  ordered.reserve(prop_items.size() + spec_items.size() + kry_items.size());
  for (auto& c : prop_items) ordered.push_back(c.second);
  for (auto& c : spec_items) ordered.push_back(c.second);
  for (auto& c : kry_items) ordered.push_back(c.second);
  pl->h_items = ordered;
  // diagnostics: a smaller cap on the eigensolver's secular iterations (the propagator-path
  // fallback for orientations it leaves unconverged is tested with it)
  if (const char* e = getenv("SPINACH_DC_MAX_ITER")) pl->dc_max_iter = std::max(0, atoi(e));
  pl->n_prop = (int64_t)prop_items.size();
  pl->n_spec = (int64_t)spec_items.size();
  pl->n_spec_w = pl->n_spec_g = 0;
  for (auto& c : spec_items) {
    const int k = spec_class(c.second);
    pl->n_spec_w += (k > 0);
    pl->n_spec_g += (k > 1);
  }
  pl->spec_nm = pl->n_spec_g > 0 ? spk::kTriGNP : (pl->n_spec_w > 0 ? spk::kTriWNP : spk::kTriNP);
  pl->n_kry = (int64_t)kry_items.size();
  pl->n_items = pl->n_prop + pl->n_spec + pl->n_kry;
  pl->total_or = or_base;

  hipStream_t st = ctx->stream;
  upload(pl->d_sys, descs, st);
  upload(pl->d_items, ordered, st);
  upload(pl->d_row, row, st);
  upload(pl->d_col, col, st);
  upload(pl->d_tpos, tpos, st);
  upload(pl->d_vals, vals, st);
  upload(pl->d_slots, slots, st);
  upload(pl->d_rho, rho, st);
  upload(pl->d_coil, coil, st);
  upload(pl->d_al, al, st);
  upload(pl->d_be, be, st);
  upload(pl->d_ga, ga, st);
  upload(pl->d_w, w, st);
  upload(pl->d_sdesc, sdescs, st);
  upload(pl->d_sell_start, sell_start, st);
  upload(pl->d_sell_col, sell_col, st);
  upload(pl->d_sell_pos, sell_pos, st);
  upload(pl->d_sell_perm, sell_perm, st);
  if (!rvals.empty()) upload(pl->d_rvals, rvals, st);
  if (!rdiag.empty()) upload(pl->d_rdiag, rdiag, st);
  pl->rho_or_stride = nmax;
  pl->step_vec_words = 8 * ((nmax + spk::kSell - 1) / spk::kSell * spk::kSell);
  pl->d_nsub.alloc((size_t)std::max<int64_t>(1, or_base) * sizeof(int32_t));
  HIPCHK(hipMemsetAsync(pl->d_nsub.p, 0, pl->d_nsub.bytes, st));  // only stepped orientations write it
  pl->d_terms.alloc((size_t)std::max<int64_t>(1, or_base) * sizeof(int32_t));
  HIPCHK(hipMemsetAsync(pl->d_terms.p, 0, pl->d_terms.bytes, st));
  pl->d_queue.alloc(spk::kQueues * sizeof(unsigned int));
  pl->d_per.alloc((size_t)pl->total_or * npts * 2 * sizeof(double));
  pl->d_fid.alloc((size_t)pl->n_systems * npts * 2 * sizeof(double));
  pl->d_sq.alloc((size_t)std::max<int64_t>(1, pl->total_or) * sizeof(int32_t));
  HIPCHK(hipMemsetAsync(pl->d_sq.p, 0, pl->d_sq.bytes, st));

  if (pl->n_prop > 0) {
    // persistent grid: resident workgroups, bounded by the item count
    int per_cu = 0;
    if (mode == SP_MODE_FP64)
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spk::k_powder<double, false>, spk::kThreads, 0));
    else
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spk::k_powder<float, false>, spk::kThreads, 0));
    // the occupancy query is conservative for this kernel; at least two
    // workgroups per CU hide the LDS-staging latency (override:
    // SPINACH_WG_PER_CU). Workgroups beyond residency find the queue drained.
    per_cu = std::max(2, per_cu);
    if (const char* e = getenv("SPINACH_WG_PER_CU")) per_cu = std::max(1, atoi(e));
    configure_prop(ctx, pl.get(), pl->log2b, per_cu);
  }
  if (pl->n_spec > 0) {
    // two 256-thread workgroups per CU, each holding the lower triangle of a 192 x 192
    // matrix in registers (one 1024-thread workgroup per CU for larger systems); three
    // synthesis workgroups per CU
    configure_spec(ctx, pl.get(), 2, 3);
    pl->d_spec.alloc((size_t)pl->n_spec * 6 * pl->spec_nm * sizeof(double));
    if (pl->n_spec_g > 0)
      pl->d_tri_work.alloc((size_t)std::min<int64_t>(ctx->n_cu, pl->n_spec_w) * spk::kTriGWork * sizeof(double));
    pl->d_status.alloc((size_t)pl->n_spec * sizeof(int32_t));
  }
  if (pl->n_kry > 0) setup_step(ctx, pl.get());
  setup_sum(pl.get(), st);
  HIPCHK(hipStreamSynchronize(st));
  apply_cached_tuning(pl.get());
  return pl.release();
}

template <typename R>
void launch(sp_plan* pl) {
  spk::KParams<R> kp{};
  kp.sys = pl->d_sys.as<spk::SysDesc>();
  kp.items = pl->d_items.as<spk::Item>();
  kp.n_items = pl->n_prop;
  kp.queue = pl->d_queue.as<unsigned int>();
  kp.prow = pl->d_row.as<int32_t>();
  kp.pcol = pl->d_col.as<int32_t>();
  kp.ptpos = pl->d_tpos.as<int32_t>();
  kp.vals = pl->d_vals.as<double>();
  kp.slots = pl->d_slots.as<int32_t>();
  kp.rho0 = pl->d_rho.as<double>();
  kp.rho_or = pl->use_rho_or() ? pl->d_rho_or.as<double>() : nullptr;
  kp.rho_or_stride = pl->rho_or_stride;
  kp.coil = pl->d_coil.as<double>();
  kp.alphas = pl->d_al.as<double>();
  kp.betas = pl->d_be.as<double>();
  kp.gammas = pl->d_ga.as<double>();
  kp.dt = pl->dt;
  kp.nsteps = pl->nsteps;
  kp.block = pl->block;
  kp.nblk = pl->nblk;
  kp.log2b = pl->log2b;
  kp.taylor_k = pl->taylor_k;
  kp.ps_q = pl->ps_q;
  kp.theta = pl->theta;
  kp.work = pl->d_work.as<R>();
  kp.work_stride = pl->work_stride;
  kp.np_max = pl->np_prop;
  kp.per_or = pl->d_per.as<double>();
  kp.sq_out = pl->d_sq.as<int32_t>();
  kp.rvals = pl->d_rvals.as<double>();
  kp.rdiag = pl->d_rdiag.as<double>();
  // complex-generator items first (queue 0), then real-generator items (queue 3)
  const int64_t n_cplx = pl->n_prop_complex, n_real = pl->n_prop - pl->n_prop_complex;
  if (n_cplx > 0) {
    kp.n_items = n_cplx;
    const int g = (int)std::min<int64_t>(pl->grid, n_cplx);
    hipLaunchKernelGGL((spk::k_powder<R, false>), dim3(g), dim3(spk::kThreads), 0, pl->ctx->stream, kp);
    HIPCHK(hipGetLastError());
  }
  if (n_real > 0) {
    kp.items = pl->d_items.as<spk::Item>() + n_cplx;
    kp.n_items = n_real;
    kp.queue = pl->d_queue.as<unsigned int>() + 3;
    const int g = (int)std::min<int64_t>(pl->grid, n_real);
    hipLaunchKernelGGL((spk::k_powder<R, true>), dim3(g), dim3(spk::kThreads), 0, pl->ctx->stream, kp);
    HIPCHK(hipGetLastError());
  }
}

spk::SpecParams spec_params(sp_plan* pl) {
  spk::SpecParams sp{};
  sp.sys = pl->d_sys.as<spk::SysDesc>();
  sp.items = pl->d_items.as<spk::Item>() + pl->n_prop;
  sp.n_items = pl->n_spec;
  sp.queue = pl->d_queue.as<unsigned int>() + 1;
  sp.prow = pl->d_row.as<int32_t>();
  sp.pcol = pl->d_col.as<int32_t>();
  sp.ptpos = pl->d_tpos.as<int32_t>();
  sp.vals = pl->d_vals.as<double>();
  sp.slots = pl->d_slots.as<int32_t>();
  sp.rho0 = pl->d_rho.as<double>();
  sp.rho_or = pl->use_rho_or() ? pl->d_rho_or.as<double>() : nullptr;
  sp.rho_or_stride = pl->rho_or_stride;
  sp.coil = pl->d_coil.as<double>();
  sp.alphas = pl->d_al.as<double>();
  sp.betas = pl->d_be.as<double>();
  sp.gammas = pl->d_ga.as<double>();
  sp.dt = pl->dt;
  sp.nsteps = pl->nsteps;
  sp.spec = pl->d_spec.as<double>();
  sp.spec_nm = pl->spec_nm;
  sp.tri_work = pl->n_spec_g > 0 ? pl->d_tri_work.as<double>() : nullptr;
  sp.tri_work_stride = spk::kTriGWork;
  sp.status = pl->d_status.as<int32_t>();
  sp.per_or = pl->d_per.as<double>();
  sp.dc_max_iter = pl->dc_max_iter;
  return sp;
}

template <bool REAL, bool VL, int RCS = 0>
void launch_step_v(sp_plan* pl, const spk::StepParams& sp, size_t lds, int grid) {
  hipLaunchKernelGGL((spk::k_step<REAL, VL, RCS>), dim3(grid), dim3(spk::kStepThreads), lds, pl->ctx->stream, sp);
  HIPCHK(hipGetLastError());
}

template <bool REAL>
void launch_step_kind(sp_plan* pl, const spk::StepParams& sp, size_t lds, int grid, bool vl) {
  if (vl && pl->step_rcs == 1)
    launch_step_v<REAL, true, 1>(pl, sp, lds, grid);
  else if (vl && pl->step_rcs == 2)
    launch_step_v<REAL, true, 2>(pl, sp, lds, grid);
  else if (vl)
    launch_step_v<REAL, true>(pl, sp, lds, grid);
  else
    launch_step_v<REAL, false>(pl, sp, lds, grid);
}

// n stepping items from items, on queue head q: the observable (final_state
// false, from rho0 or the dead-time states) or the state after nsteps steps of
// dt written to rho_out (final_state true, from rho0).
void launch_step_items(sp_plan* pl, bool real, const spk::Item* items, int64_t n, int q, bool final_state, double dt,
                       int64_t nsteps, double* rho_out, const double* rho_in = nullptr) {
  spk::StepParams sp{};
  sp.sys = pl->d_sys.as<spk::SysDesc>();
  sp.sdesc = pl->d_sdesc.as<spk::StepDesc>();
  sp.items = items;
  sp.n_items = n;
  sp.queue = pl->d_queue.as<unsigned int>() + q;
  sp.ptpos = pl->d_tpos.as<int32_t>();
  sp.vals = pl->d_vals.as<double>();
  sp.slots = pl->d_slots.as<int32_t>();
  sp.sell_start = pl->d_sell_start.as<int32_t>();
  sp.sell_col = pl->d_sell_col.as<int32_t>();
  sp.sell_pos = pl->d_sell_pos.as<int32_t>();
  sp.sell_perm = pl->d_sell_perm.as<int32_t>();
  sp.rho0 = pl->d_rho.as<double>();
  // the acquisition starts from the per-orientation states; a final-state pass from rho_in
  sp.rho_or = final_state ? rho_in : (pl->use_rho_or() ? pl->d_rho_or.as<double>() : nullptr);
  sp.rho_or_stride = pl->rho_or_stride;
  sp.coil = pl->d_coil.as<double>();
  sp.alphas = pl->d_al.as<double>();
  sp.betas = pl->d_be.as<double>();
  sp.gammas = pl->d_ga.as<double>();
  sp.dt = dt;
  sp.nsteps = nsteps;
  sp.final_state = final_state ? 1 : 0;
  sp.vec_words = pl->step_vec_words;
  sp.vec_global = pl->step_vec_global;
  sp.gvec_off = pl->step_gvec_off;
  sp.work = pl->d_step_work.as<double>();
  sp.work_stride = (int64_t)(pl->d_step_work.bytes / sizeof(double) / std::max(1, pl->grid_step));
  sp.per_or = pl->d_per.as<double>();
  sp.rho_out = rho_out;
  sp.nsub_out = final_state ? pl->d_nsub_fs.as<int32_t>() : pl->d_nsub.as<int32_t>();
  sp.terms_out = final_state ? nullptr : pl->d_terms.as<int32_t>();
  sp.prow = pl->d_row.as<int32_t>();
  sp.rvals = pl->d_rvals.as<double>();
  sp.rdiag = pl->d_rdiag.as<double>();
  const int g = (int)std::min<int64_t>(pl->grid_step, n);
  const int r = real ? 1 : 0;
  if (real)
    launch_step_kind<true>(pl, sp, pl->step_lds[r], g, pl->step_vl[r]);
  else
    launch_step_kind<false>(pl, sp, pl->step_lds[r], g, pl->step_vl[r]);
}

// The stepping items of a plan: complex generators, then real ones.
void launch_krylov(sp_plan* pl) {
  const spk::Item* items = pl->d_items.as<spk::Item>() + pl->n_prop + pl->n_spec;
  const int64_t n_c = pl->n_kry_complex, n_r = pl->n_kry - pl->n_kry_complex;
  if (n_c > 0) launch_step_items(pl, false, items, n_c, 4, false, pl->dt, pl->nsteps, nullptr);
  if (n_r > 0) launch_step_items(pl, true, items + n_c, n_r, 5, false, pl->dt, pl->nsteps, nullptr);
}

// The spectral items from `first` on, n of them (item, spectrum and status pointers offset).
spk::SpecParams spec_params_range(sp_plan* pl, int64_t first, int64_t n) {
  spk::SpecParams sp = spec_params(pl);
  sp.items += first;
  sp.n_items = n;
  sp.spec += (size_t)first * 6 * pl->spec_nm;
  sp.status += first;
  return sp;
}

// The three spectral kernels; ev[0..2] are recorded after each of them. Items above kTriNP
// (the first n_spec_w) take the 1024-thread tridiagonalisation (queue head 8), with its
// global-memory prologue for the first n_spec_g (above kTriWNP); the eigensolver is sized per
// class; the synthesis runs over every spectral item.
void launch_spectral(sp_plan* pl, hipEvent_t* ev) {  // ev == nullptr: no stage events (graph recording)
  hipStream_t st = pl->ctx->stream;
  const int64_t nw = pl->n_spec_w, nn = pl->n_spec - pl->n_spec_w;
  if (nw > 0) {
    spk::SpecParams tw = spec_params_range(pl, 0, nw);
    tw.queue = pl->d_queue.as<unsigned int>() + 8;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)pl->ctx->n_cu, nw));
    hipLaunchKernelGGL(spk::k_tridiag_w, dim3(g), dim3(spk::kTriWThreads), 0, st, tw);
    HIPCHK(hipGetLastError());
  }
  if (nn > 0) {
    spk::SpecParams tri = spec_params_range(pl, nw, nn);
    hipLaunchKernelGGL(spk::k_tridiag, dim3(std::max<int64_t>(1, std::min<int64_t>(pl->grid_tri, nn))),
                       dim3(spk::kTriThreads), 0, st, tri);
    HIPCHK(hipGetLastError());
  }
  if (ev) HIPCHK(hipEventRecord(ev[0], st));
  const int64_t ng = pl->n_spec_g;
  if (ng > 0) {
    spk::SpecParams tg = spec_params_range(pl, 0, ng);
    hipLaunchKernelGGL(spk::k_dceig<spk::kTriGNP>, dim3((unsigned)ng), dim3(spk::DcSharedT<spk::kTriGNP>::kThr), 0,
                       st, tg);
    HIPCHK(hipGetLastError());
  }
  if (nw > ng) {
    spk::SpecParams tw = spec_params_range(pl, ng, nw - ng);
    hipLaunchKernelGGL(spk::k_dceig<spk::kTriWNP>, dim3((unsigned)(nw - ng)), dim3(spk::DcSharedT<spk::kTriWNP>::kThr),
                       0, st, tw);
    HIPCHK(hipGetLastError());
  }
  if (nn > 0) {
    spk::SpecParams tri = spec_params_range(pl, nw, nn);
    hipLaunchKernelGGL(spk::k_dceig<spk::kTriNP>, dim3((unsigned)nn), dim3(spk::DcSharedT<spk::kTriNP>::kThr), 0, st,
                       tri);
    HIPCHK(hipGetLastError());
  }
  if (ev) HIPCHK(hipEventRecord(ev[1], st));
  spk::SpecParams syn = spec_params(pl);
  if (pl->spec_nm == spk::kTriGNP)
    hipLaunchKernelGGL(spk::k_synth<spk::kTriGNP>, dim3(pl->grid_syn), dim3(spk::kThreads), 0, st, syn);
  else if (pl->spec_nm == spk::kTriWNP)
    hipLaunchKernelGGL(spk::k_synth<spk::kTriWNP>, dim3(pl->grid_syn), dim3(spk::kThreads), 0, st, syn);
  else
    hipLaunchKernelGGL(spk::k_synth<spk::kTriNP>, dim3(pl->grid_syn), dim3(spk::kThreads), 0, st, syn);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord(ev[2], st));
}

// Wigner matrices (rank 1 then rank 2, interleaved complex) at Euler angles.
std::vector<double> wigner_pair(double a, double b, double g) {
  double cre[34], cim[34];
  spk::wigner_coeffs(a, b, g, cre, cim);
  std::vector<double> v(68);
  for (int t = 0; t < 34; t++) {
    v[2 * t] = cre[t];
    v[2 * t + 1] = cim[t];
  }
  return v;
}

sp_plan* build_mas_plan(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* al, const double* be,
                        const double* ga, const double* w, int64_t n_or, const double* rho0, const double* coil,
                        const sp_rotor* rot, int64_t nsteps, int mode) {
  if (!I || !Q) throw SpError(SP_EINVAL, "I and Q must not be NULL");
  if (!al || !be || !ga || !w || n_or <= 0) throw SpError(SP_EINVAL, "invalid grid");
  if (!rho0 || !coil) throw SpError(SP_EINVAL, "rho0 and coil must not be NULL");
  if (!rot) throw SpError(SP_EINVAL, "rotor must not be NULL");
  if (!(rot->rate_hz > 0) || !std::isfinite(rot->rate_hz)) throw SpError(SP_EINVAL, "rotor rate must be positive");
  if (rot->ticks < 1 || rot->ticks > 4096) throw SpError(SP_EINVAL, "rotor ticks must be in 1..4096");
  if (rot->frame != SP_MAS_FRAME_MAGNET && rot->frame != SP_MAS_FRAME_ROTOR)
    throw SpError(SP_EINVAL, "unknown MAS frame");
  double ax = rot->axis[0], ay = rot->axis[1], az = rot->axis[2];
  double an = std::sqrt(ax * ax + ay * ay + az * az);
  if (!(an > 0) || !std::isfinite(an)) throw SpError(SP_EINVAL, "rotor axis must be a non-zero finite vector");
  if (nsteps < 0 || nsteps > (int64_t)1 << 30) throw SpError(SP_EINVAL, "invalid nsteps");
  const bool graph = (mode & SP_MODE_GRAPH) != 0;
  mode &= ~(SP_MODE_COMPLEX | SP_MODE_PROPAGATOR | SP_MODE_GRAPH);  // the rotor path is always complex propagation
  if (mode != SP_MODE_FP64 && mode != SP_MODE_FP32) throw SpError(SP_EINVAL, "unknown precision mode");
  const int64_t N = I->n;
  if (N <= 0 || N > 1024) throw SpError(SP_ERANGE, "Hilbert dimension must be in 1..1024");
  std::vector<Triplets> mats(1 + SP_NQ);
  mats[0] = read_csc(I, N, "I");
  for (int q = 0; q < SP_NQ; q++) mats[1 + q] = read_csc(&Q[q], N, "Q slot");
  for (int64_t k = 0; k < n_or; k++)
    if (!std::isfinite(al[k]) || !std::isfinite(be[k]) || !std::isfinite(ga[k]) || !std::isfinite(w[k]))
      throw SpError(SP_EINVAL, "non-finite grid entry");
  for (int64_t i = 0; i < 2 * N * N; i++)
    if (!std::isfinite(rho0[i]) || !std::isfinite(coil[i])) throw SpError(SP_EINVAL, "non-finite state matrix");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<sp_plan> pl(new sp_plan);
  pl->ctx = ctx;
  pl->mas = true;
  pl->mode = mode;
  pl->graph = graph;
  pl->nsteps = nsteps;
  pl->n_systems = 1;
  pl->ticks = rot->ticks;
  pl->dt = 1.0 / (rot->rate_hz * rot->ticks);
  const int64_t npts = nsteps + 1;
  pl->n_rho = (int32_t)((npts + rot->ticks - 1) / rot->ticks);
  const int32_t np = (int32_t)std::max<int64_t>(16, (N + 15) / 16 * 16);
  pl->mas_np = np;
  pl->np_max = np;
  pl->theta = 1.0;
  if (mode == SP_MODE_FP64) {
    pl->taylor_k = 18;
    pl->ps_q = 4;
  } else {
    pl->taylor_k = 10;
    pl->ps_q = 3;
  }
  // union pattern of I and Q (no reduction in Hilbert space)
  std::vector<std::pair<int64_t, int64_t>> ent;
  for (auto& m : mats)
    for (size_t q = 0; q < m.row.size(); q++) ent.emplace_back(m.row[q], m.col[q]);
  std::sort(ent.begin(), ent.end());
  ent.erase(std::unique(ent.begin(), ent.end()), ent.end());
  const int32_t nnz = (int32_t)ent.size();
  std::vector<int32_t> row(nnz), col(nnz), slots;
  for (int32_t e = 0; e < nnz; e++) {
    row[e] = (int32_t)ent[e].first;
    col[e] = (int32_t)ent[e].second;
  }
  auto find = [&](int64_t r, int64_t c) {
    return (int64_t)(std::lower_bound(ent.begin(), ent.end(), std::make_pair(r, c)) - ent.begin());
  };
  std::vector<double> vals, tmp;
  auto scatter = [&](const Triplets& m) {
    tmp.assign(2 * (size_t)nnz, 0.0);
    for (size_t q = 0; q < m.row.size(); q++) {
      int64_t e = find(m.row[q], m.col[q]);
      tmp[2 * e] += m.re[q];
      tmp[2 * e + 1] += m.im[q];
    }
    return !m.row.empty();
  };
  scatter(mats[0]);
  vals = tmp;
  for (int q = 0; q < SP_NQ; q++)
    if (scatter(mats[1 + q])) {
      slots.push_back(q);
      vals.insert(vals.end(), tmp.begin(), tmp.end());
    }
  pl->mas_nnz = nnz;
  pl->mas_nq = (int32_t)slots.size();
  // dense states: MATLAB column-major N x N -> padded row-major np x np
  std::vector<double> r0(2 * (size_t)np * np, 0.0), c0(2 * (size_t)np * np, 0.0);
  for (int64_t c = 0; c < N; c++)
    for (int64_t r = 0; r < N; r++) {
      size_t src = 2 * (size_t)(c * N + r), dst = 2 * (size_t)(r * np + c);
      r0[dst] = rho0[src];
      r0[dst + 1] = rho0[src + 1];
      c0[dst] = coil[src];
      c0[dst + 1] = coil[src + 1];
    }
  // rotor frame matrices (rotor_stack.m:88-105): cart2sph of the axis
  double rphi = std::atan2(ay, ax);
  double rtheta = M_PI / 2 - std::atan2(az, std::sqrt(ax * ax + ay * ay));
  std::vector<double> rot2lab = wigner_pair(rphi, rtheta, 0.0);
  std::vector<double> lab2rot(68, 0.0);
  if (rot->frame == SP_MAS_FRAME_MAGNET) {
    lab2rot = wigner_pair(0.0, -rtheta, -rphi);
  } else {
    for (int k = 0; k < 3; k++) lab2rot[2 * (k * 3 + k)] = 1.0;
    for (int k = 0; k < 5; k++) lab2rot[2 * (9 + k * 5 + k)] = 1.0;
  }
  hipStream_t st = ctx->stream;
  // invariant subspaces of every H(phi): connected components of the union pattern. When none
  // exceeds 16 states (and in FP64), pack them into tiles of 16, first-fit decreasing, and
  // propagate tile by tile (spk::k_mas_tiled); SPINACH_MAS_DENSE=1 keeps the dense kernel.
  {
    std::vector<int64_t> par(N);
    for (int64_t i = 0; i < N; i++) par[i] = i;
    auto root = [&](int64_t x) {
      while (par[x] != x) x = par[x] = par[par[x]];
      return x;
    };
    for (int32_t e = 0; e < nnz; e++) {
      const int64_t a = root(row[e]), b = root(col[e]);
      if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
    std::vector<std::vector<int64_t>> comps;
    std::vector<int64_t> cid(N, -1);
    for (int64_t i = 0; i < N; i++) {
      const int64_t rt = root(i);
      if (cid[rt] < 0) {
        cid[rt] = (int64_t)comps.size();
        comps.emplace_back();
      }
      comps[cid[rt]].push_back(i);
    }
    size_t biggest = 0;
    for (auto& c : comps) biggest = std::max(biggest, c.size());
    const char* ed = getenv("SPINACH_MAS_DENSE");
    if (mode == SP_MODE_FP64 && biggest <= (size_t)spk::kTileN && !(ed && atoi(ed))) {
      std::vector<size_t> order(comps.size());
      for (size_t k = 0; k < order.size(); k++) order[k] = k;
      std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return comps[a].size() > comps[b].size(); });
      std::vector<int> used;                       // states placed in each tile
      std::vector<int64_t> tile_of(N), local(N);   // tile and position of each state
      for (size_t k : order) {
        const int sz = (int)comps[k].size();
        size_t t = 0;
        while (t < used.size() && used[t] + sz > spk::kTileN) t++;
        if (t == used.size()) used.push_back(0);
        for (int64_t i : comps[k]) {
          tile_of[i] = (int64_t)t;
          local[i] = used[t]++;
        }
      }
      const int ntile = (int)used.size();
      if (ntile <= spk::kMasMaxTiles) {
        // pattern entries by tile, with tile-local positions; values in the same order
        std::vector<int32_t> eorder(nnz);
        for (int32_t e = 0; e < nnz; e++) eorder[e] = e;
        std::stable_sort(eorder.begin(), eorder.end(),
                         [&](int32_t a, int32_t b) { return tile_of[row[a]] < tile_of[row[b]]; });
        std::vector<int32_t> epos(nnz), toff(ntile + 1, 0);
        const int nplanes = 1 + (int)slots.size();
        std::vector<double> tvals(vals.size());
        for (int32_t k = 0; k < nnz; k++) {
          const int32_t e = eorder[k];
          epos[k] = (int32_t)(local[row[e]] * spk::kTileN + local[col[e]]);
          toff[tile_of[row[e]] + 1]++;
          for (int pln = 0; pln < nplanes; pln++) {
            tvals[2 * ((size_t)pln * nnz + k)] = vals[2 * ((size_t)pln * nnz + e)];
            tvals[2 * ((size_t)pln * nnz + k) + 1] = vals[2 * ((size_t)pln * nnz + e) + 1];
          }
        }
        for (int t = 0; t < ntile; t++) toff[t + 1] += toff[t];
        std::vector<int32_t> trow(nnz), tcol(nnz);  // row and col follow the values' order
        for (int32_t k = 0; k < nnz; k++) {
          trow[k] = row[eorder[k]];
          tcol[k] = col[eorder[k]];
        }
        // tile pairs where both coil and rho0 have entries (MATLAB column-major N x N input)
        std::vector<double> cblk((size_t)ntile * ntile * 512, 0.0), rblk((size_t)ntile * ntile * 512, 0.0);
        std::vector<char> cnz((size_t)ntile * ntile, 0), rnz((size_t)ntile * ntile, 0);
        for (int64_t c = 0; c < N; c++)
          for (int64_t r = 0; r < N; r++) {
            const size_t src = 2 * (size_t)(c * N + r);
            const size_t pr = (size_t)tile_of[r] * ntile + tile_of[c];
            const size_t dst = pr * 512 + 2 * (size_t)(local[r] * spk::kTileN + local[c]);
            if (coil[src] != 0.0 || coil[src + 1] != 0.0) {
              cnz[pr] = 1;
              cblk[dst] = coil[src];
              cblk[dst + 1] = coil[src + 1];
            }
            if (rho0[src] != 0.0 || rho0[src + 1] != 0.0) {
              rnz[pr] = 1;
              rblk[dst] = rho0[src];
              rblk[dst + 1] = rho0[src + 1];
            }
          }
        std::vector<int32_t> ps, pt;
        std::vector<double> ct, rt;
        for (int a = 0; a < ntile; a++)
          for (int b = 0; b < ntile; b++) {
            const size_t pr = (size_t)a * ntile + b;
            if (cnz[pr] && rnz[pr]) {
              ps.push_back(a);
              pt.push_back(b);
              ct.insert(ct.end(), cblk.begin() + pr * 512, cblk.begin() + (pr + 1) * 512);
              rt.insert(rt.end(), rblk.begin() + pr * 512, rblk.begin() + (pr + 1) * 512);
            }
          }
        if (ps.empty()) {  // nothing detected: keep one pair so the kernel's shapes stay valid
          ps.push_back(0);
          pt.push_back(0);
          ct.assign(512, 0.0);
          rt.assign(512, 0.0);
        }
        // Within a pair's tiles, C_j = U_j' coil U_j and rho_k = U^k rho0 U'^k are nonzero only on
        // the component blocks (a, b) where coil (resp. rho0) has entries, since U_j is block
        // diagonal over the components. The detection needs the elements where both can be
        // nonzero: emap numbers them (pair, tile-local position -> compact index, -1 elsewhere).
        const size_t ncomp = comps.size();
        std::vector<char> ccomp(ncomp * ncomp, 0), rcomp(ncomp * ncomp, 0);
        for (int64_t c = 0; c < N; c++)
          for (int64_t r = 0; r < N; r++) {
            const size_t src = 2 * (size_t)(c * N + r);
            const size_t cp = (size_t)cid[root(r)] * ncomp + cid[root(c)];
            if (coil[src] != 0.0 || coil[src + 1] != 0.0) ccomp[cp] = 1;
            if (rho0[src] != 0.0 || rho0[src + 1] != 0.0) rcomp[cp] = 1;
          }
        std::vector<int64_t> state_at((size_t)ntile * spk::kTileN, -1);
        for (int64_t i = 0; i < N; i++) state_at[(size_t)tile_of[i] * spk::kTileN + local[i]] = i;
        std::vector<int32_t> emap(ps.size() * 256, -1);
        int32_t ndet = 0;
        for (size_t pp = 0; pp < ps.size(); pp++)
          for (int r = 0; r < spk::kTileN; r++)
            for (int c = 0; c < spk::kTileN; c++) {
              const int64_t i = state_at[(size_t)ps[pp] * spk::kTileN + r];
              const int64_t k = state_at[(size_t)pt[pp] * spk::kTileN + c];
              if (i < 0 || k < 0) continue;
              const size_t cp = (size_t)cid[root(i)] * ncomp + cid[root(k)];
              if (ccomp[cp] && rcomp[cp]) emap[pp * 256 + r * spk::kTileN + c] = ndet++;
            }
        pl->mas_ldet = std::max<int32_t>(16, (ndet + 15) / 16 * 16);
        pl->mas_coil_real = true;
        for (size_t k = 1; k < ct.size(); k += 2)
          if (ct[k] != 0.0) pl->mas_coil_real = false;
        upload(pl->d_emap, emap, st);
        pl->mas_tiled = true;
        pl->np_max = spk::kTileN;  // reported as the padded dimension: every product is 16 x 16
        pl->mas_ntile = ntile;
        pl->mas_npair = (int32_t)ps.size();
        upload(pl->d_epos, epos, st);
        upload(pl->d_toff, toff, st);
        upload(pl->d_pair_s, ps, st);
        upload(pl->d_pair_t, pt, st);
        upload(pl->d_coil_t, ct, st);
        upload(pl->d_rho0_t, rt, st);
        vals.swap(tvals);
        row.swap(trow);
        col.swap(tcol);
      }
    }
  }
  upload(pl->d_row, row, st);
  upload(pl->d_col, col, st);
  upload(pl->d_vals, vals, st);
  upload(pl->d_slots, slots, st);
  upload(pl->d_rho, r0, st);
  upload(pl->d_coil, c0, st);
  upload(pl->d_al, std::vector<double>(al, al + n_or), st);
  upload(pl->d_be, std::vector<double>(be, be + n_or), st);
  upload(pl->d_ga, std::vector<double>(ga, ga + n_or), st);
  upload(pl->d_w, std::vector<double>(w, w + n_or), st);
  upload(pl->d_rot2lab, rot2lab, st);
  upload(pl->d_lab2rot, lab2rot, st);
  pl->d_queue.alloc(spk::kQueues * sizeof(unsigned int));
  pl->n_items = n_or;
  pl->total_or = n_or;
  pl->or_off.push_back(0);
  pl->n_or.push_back(n_or);
  pl->full_dim = N;
  pl->red_dim = N;
  pl->red_nnz = nnz;
  int per_cu = pl->mas_tiled ? 3 : 2;  // k_mas_tiled: 168 VGPRs and 52.5 KB of LDS, three per CU
  if (const char* e = getenv("SPINACH_WG_PER_CU")) per_cu = std::max(1, atoi(e));
  pl->per_cu = per_cu;
  pl->grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * per_cu, n_or));
  const size_t esz = (mode == SP_MODE_FP64) ? sizeof(double) : sizeof(float);
  const int64_t mat = 2 * (int64_t)np * np;
  if (pl->mas_tiled) {
    // per workgroup: C (ticks x ldet) and rho (n_rho x ldet) on the detected elements, planar
    // complex FP64, and the five Fourier components of the generator per pattern entry
    const int64_t Lt = pl->mas_ldet;
    pl->work_stride = 2 * Lt * ((int64_t)pl->ticks + pl->n_rho) + (int64_t)nnz * 10;
  } else {
    pl->work_stride = 10 * mat + mat * (int64_t)pl->ticks + mat * (int64_t)pl->n_rho;
  }
  pl->work_stride = (pl->work_stride + 63) / 64 * 64;
  pl->d_work.alloc((size_t)pl->grid * pl->work_stride * esz);
  // the padding columns of the detected elements (ldet - elements) are never written: zero once
  if (pl->mas_tiled) HIPCHK(hipMemsetAsync(pl->d_work.p, 0, (size_t)pl->grid * pl->work_stride * esz, st));
  pl->d_per.alloc((size_t)n_or * npts * 2 * sizeof(double));
  pl->d_fid.alloc((size_t)npts * 2 * sizeof(double));
  pl->d_sq.alloc((size_t)n_or * sizeof(int32_t));
  if (pl->mas_tiled) pl->d_nprod.alloc((size_t)n_or * sizeof(int32_t));
  setup_sum(pl.get(), st);
  HIPCHK(hipStreamSynchronize(st));
  return pl.release();
}

template <typename R>
void launch_mas(sp_plan* pl) {
  spk::MasParams<R> kp{};
  kp.prow = pl->d_row.as<int32_t>();
  kp.pcol = pl->d_col.as<int32_t>();
  kp.nnz = pl->mas_nnz;
  kp.nq = pl->mas_nq;
  kp.vals = pl->d_vals.as<double>();
  kp.slots = pl->d_slots.as<int32_t>();
  kp.rho0 = pl->d_rho.as<double>();
  kp.coil = pl->d_coil.as<double>();
  kp.alphas = pl->d_al.as<double>();
  kp.betas = pl->d_be.as<double>();
  kp.gammas = pl->d_ga.as<double>();
  kp.n_or = pl->total_or;
  kp.rot2lab = pl->d_rot2lab.as<double>();
  kp.lab2rot = pl->d_lab2rot.as<double>();
  kp.dt = pl->dt;
  kp.ticks = pl->ticks;
  kp.n_rho = pl->n_rho;
  kp.nsteps = pl->nsteps;
  kp.np = pl->mas_np;
  kp.taylor_k = pl->taylor_k;
  kp.ps_q = pl->ps_q;
  kp.theta = pl->theta;
  kp.work = pl->d_work.as<R>();
  kp.work_stride = pl->work_stride;
  kp.queue = pl->d_queue.as<unsigned int>();
  kp.per_or = pl->d_per.as<double>();
  kp.sq_out = pl->d_sq.as<int32_t>();
  hipLaunchKernelGGL(spk::k_mas<R>, dim3(pl->grid), dim3(spk::kThreads), 0, pl->ctx->stream, kp);
  HIPCHK(hipGetLastError());
}

void launch_mas_tiled(sp_plan* pl) {
  spk::MasTParams kp{};
  kp.epos = pl->d_epos.as<int32_t>();
  kp.toff = pl->d_toff.as<int32_t>();
  kp.ntile = pl->mas_ntile;
  kp.nnz = pl->mas_nnz;
  kp.nq = pl->mas_nq;
  kp.vals = pl->d_vals.as<double>();
  kp.slots = pl->d_slots.as<int32_t>();
  kp.npair = pl->mas_npair;
  kp.coil_real = pl->mas_coil_real ? 1 : 0;
  kp.pair_s = pl->d_pair_s.as<int32_t>();
  kp.pair_t = pl->d_pair_t.as<int32_t>();
  kp.coil_t = pl->d_coil_t.as<double>();
  kp.rho0_t = pl->d_rho0_t.as<double>();
  kp.emap = pl->d_emap.as<int32_t>();
  kp.ldet = pl->mas_ldet;
  kp.alphas = pl->d_al.as<double>();
  kp.betas = pl->d_be.as<double>();
  kp.gammas = pl->d_ga.as<double>();
  kp.n_or = pl->total_or;
  kp.rot2lab = pl->d_rot2lab.as<double>();
  kp.lab2rot = pl->d_lab2rot.as<double>();
  kp.dt = pl->dt;
  kp.ticks = pl->ticks;
  kp.n_rho = pl->n_rho;
  kp.nsteps = pl->nsteps;
  kp.taylor_k = pl->taylor_k;
  kp.ps_q = pl->ps_q;
  kp.theta = pl->theta;
  kp.work = pl->d_work.as<double>();
  kp.work_stride = pl->work_stride;
  kp.queue = pl->d_queue.as<unsigned int>();
  kp.per_or = pl->d_per.as<double>();
  kp.sq_out = pl->d_sq.as<int32_t>();
  kp.prod_out = pl->d_nprod.as<int32_t>();
  hipLaunchKernelGGL(spk::k_mas_tiled, dim3(pl->grid), dim3(spk::kThreads), 0, pl->ctx->stream, kp);
  HIPCHK(hipGetLastError());
}

// One pass on the context's stream: queue reset, propagation, orientation sum.
// With stage events (eager execution) the boundaries ev1, evs[0..3], ev2 are
// recorded; without (graph recording) nothing but the kernels is enqueued.
void enqueue_pass(sp_plan* pl, bool events) {
  sp_ctx* ctx = pl->ctx;
  hipStream_t st = ctx->stream;
  const int64_t npts = pl->nsteps + 1;
  hipLaunchKernelGGL(spk::k_reset, dim3(1), dim3(64), 0, st, pl->d_queue.as<unsigned int>());
  HIPCHK(hipGetLastError());
  if (events) HIPCHK(hipEventRecord(ctx->ev1, st));
  if (pl->mas) {
    if (pl->mas_tiled)
      launch_mas_tiled(pl);
    else if (pl->mode == SP_MODE_FP64)
      launch_mas<double>(pl);
    else
      launch_mas<float>(pl);
    if (events)
      for (auto& e : ctx->evs) HIPCHK(hipEventRecord(e, st));
  } else {
    // dead time: rho0(Omega) = step(L(Omega), rho0, dead_time) for every item (acquire.m:65-67)
    if (pl->dead_time > 0)
      launch_step_items(pl, false, pl->d_items.as<spk::Item>(), pl->n_items, 6, true, pl->dead_time, 1,
                        pl->d_rho_or.as<double>());
    if (pl->n_prop > 0) {
      if (pl->mode == SP_MODE_FP64)
        launch<double>(pl);
      else
        launch<float>(pl);
    }
    if (pl->n_kry > 0) launch_krylov(pl);
    if (events) HIPCHK(hipEventRecord(ctx->evs[0], st));
    if (pl->n_spec > 0) {
      launch_spectral(pl, events ? ctx->evs + 1 : nullptr);
    } else if (events) {
      for (int k = 1; k < 4; k++) HIPCHK(hipEventRecord(ctx->evs[k], st));
    }
  }
  if (events) HIPCHK(hipEventRecord(ctx->ev2, st));
  const unsigned nb = (unsigned)((npts + 255) / 256);
  hipLaunchKernelGGL(spk::k_chunk_sum, dim3(nb, (unsigned)pl->n_chunks), dim3(256), 0, st, pl->d_per.as<double>(),
                     pl->d_w.as<double>(), pl->d_chunks.as<spk::SumChunk>(), npts, pl->d_partial.as<double>());
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(spk::k_system_sum, dim3(nb, (unsigned)pl->n_systems), dim3(256), 0, st,
                     pl->d_partial.as<double>(), pl->d_chunk_off.as<int64_t>(), npts, pl->d_fid.as<double>());
  HIPCHK(hipGetLastError());
}

// Records the pass as a graph (once per plan); the stream is left out of
// capture mode whatever happens while recording.
void record_pass(sp_plan* pl) {
  hipStream_t st = pl->ctx->stream;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_pass(pl, false);
  } catch (...) {
    if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(st, &g));
  hipError_t ie = hipGraphInstantiate(&pl->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    pl->gexec = nullptr;
    HIPCHK(ie);
  }
}

// Spectral orientations whose eigensolution did not converge, re-run on the propagator path
// (k_powder with real products: their generators are real) into the same per-orientation
// FIDs, then the orientation sum again.
void run_fallback(sp_plan* pl, const std::vector<spk::Item>& fb) {
  sp_ctx* ctx = pl->ctx;
  hipStream_t st = ctx->stream;
  int np = 16;
  for (const auto& it : fb) np = std::max(np, pl->h_np[it.sys]);
  const int64_t npts = pl->nsteps + 1;
  const int block = std::max(1, pl->block);
  const int nblk = (int)((npts + block - 1) / block);
  int lg = 0;
  while ((1 << lg) < block) lg++;
  const int64_t mat = 2 * (int64_t)np * np;
  const int64_t stride = (6 * mat + 2 * (int64_t)(block + 16) * np + 2 * (int64_t)nblk * np + 63) / 64 * 64;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->n_cu * 2, (int64_t)fb.size()));
  DevBuf work, items;
  work.alloc((size_t)grid * stride * sizeof(double));
  upload(items, fb, st);
  spk::KParams<double> kp{};
  kp.sys = pl->d_sys.as<spk::SysDesc>();
  kp.items = items.as<spk::Item>();
  kp.n_items = (int64_t)fb.size();
  kp.queue = pl->d_queue.as<unsigned int>() + 7;
  kp.prow = pl->d_row.as<int32_t>();
  kp.pcol = pl->d_col.as<int32_t>();
  kp.ptpos = pl->d_tpos.as<int32_t>();
  kp.vals = pl->d_vals.as<double>();
  kp.slots = pl->d_slots.as<int32_t>();
  kp.rho0 = pl->d_rho.as<double>();
  kp.rho_or = pl->use_rho_or() ? pl->d_rho_or.as<double>() : nullptr;
  kp.rho_or_stride = pl->rho_or_stride;
  kp.coil = pl->d_coil.as<double>();
  kp.alphas = pl->d_al.as<double>();
  kp.betas = pl->d_be.as<double>();
  kp.gammas = pl->d_ga.as<double>();
  kp.dt = pl->dt;
  kp.nsteps = pl->nsteps;
  kp.block = block;
  kp.nblk = nblk;
  kp.log2b = lg;
  kp.taylor_k = 18;  // the FP64 Taylor degree and block length of build_plan
  kp.ps_q = 4;
  kp.theta = 1.0;
  kp.work = work.as<double>();
  kp.work_stride = stride;
  kp.np_max = np;
  kp.per_or = pl->d_per.as<double>();
  kp.sq_out = pl->d_sq.as<int32_t>();
  kp.rvals = pl->d_rvals.as<double>();
  kp.rdiag = pl->d_rdiag.as<double>();
  hipLaunchKernelGGL(spk::k_reset, dim3(1), dim3(64), 0, st, pl->d_queue.as<unsigned int>());
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL((spk::k_powder<double, true>), dim3(grid), dim3(spk::kThreads), 0, st, kp);
  HIPCHK(hipGetLastError());
  const unsigned nb = (unsigned)((npts + 255) / 256);
  hipLaunchKernelGGL(spk::k_chunk_sum, dim3(nb, (unsigned)pl->n_chunks), dim3(256), 0, st, pl->d_per.as<double>(),
                     pl->d_w.as<double>(), pl->d_chunks.as<spk::SumChunk>(), npts, pl->d_partial.as<double>());
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(spk::k_system_sum, dim3(nb, (unsigned)pl->n_systems), dim3(256), 0, st,
                     pl->d_partial.as<double>(), pl->d_chunk_off.as<int64_t>(), npts, pl->d_fid.as<double>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
}

int run(sp_plan* pl, sp_stats* stats) {
  sp_ctx* ctx = pl->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const bool graph = pl->graph;
  if (graph && !pl->gexec) record_pass(pl);
  HIPCHK(hipEventRecord(ctx->ev0, st));
  if (graph)
    HIPCHK(hipGraphLaunch(pl->gexec, st));
  else
    enqueue_pass(pl, true);
  hipEvent_t evend;
  HIPCHK(hipEventCreate(&evend));
  HIPCHK(hipEventRecord(evend, st));
  hipError_t se = hipEventSynchronize(evend);
  float ms_all = 0, ms_k = 0, ms_stage[4] = {0, 0, 0, 0};
  if (se == hipSuccess) {
    (void)hipEventElapsedTime(&ms_all, ctx->ev0, evend);
    if (graph) {
      ms_k = ms_all;
    } else {
      (void)hipEventElapsedTime(&ms_k, ctx->ev1, ctx->ev2);
      (void)hipEventElapsedTime(&ms_stage[0], ctx->ev1, ctx->evs[0]);
      for (int k = 1; k < 4; k++) (void)hipEventElapsedTime(&ms_stage[k], ctx->evs[k - 1], ctx->evs[k]);
    }
  }
  (void)hipEventDestroy(evend);
  HIPCHK(se);
  int32_t max_nsub = 0;
  int64_t step_products = 0;
  if (pl->dead_time > 0) {  // the dead-time step skips orientations whose substeps exceed the limit
    std::vector<int32_t> ns((size_t)pl->total_or);
    HIPCHK(hipMemcpy(ns.data(), pl->d_nsub_fs.p, ns.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t v : ns)
      if (v > spk::kStepMaxSub)
        throw SpError(SP_ERANGE, "either dead_time is too long, or L is too big: |L*dead_time|>1e4, check both.");
  }
  if (pl->n_kry > 0) {
    std::vector<int32_t> ns((size_t)pl->total_or), tm((size_t)pl->total_or);
    HIPCHK(hipMemcpy(ns.data(), pl->d_nsub.p, ns.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tm.data(), pl->d_terms.p, tm.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t v : ns) max_nsub = std::max(max_nsub, v);
    for (int32_t v : tm) step_products += v;
    if (max_nsub > spk::kStepMaxSub)
      throw SpError(SP_ERANGE, "either dt is too long, or L is too big: |L*dt|>1e4, check both.");
  }
  int32_t eig_failures = 0, max_sweeps = 0;
  if (pl->n_spec > 0) {
    std::vector<int32_t> stv((size_t)pl->n_spec);
    HIPCHK(hipMemcpy(stv.data(), pl->d_status.p, stv.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<spk::Item> fb;
    for (size_t k = 0; k < stv.size(); k++) {
      if (stv[k] < 0) {
        eig_failures++;
        fb.push_back(pl->h_items[(size_t)pl->n_prop + k]);
      }
      max_sweeps = std::max(max_sweeps, stv[k]);
    }
    if (!fb.empty()) run_fallback(pl, fb);
  }
  pl->eig_fallbacks = eig_failures;
  pl->executed = false;
  sp_stats s{};
  s.elapsed_s = ms_all * 1e-3;
  s.kernel_ms = ms_k;
  s.orient_steps_per_s = s.elapsed_s > 0 ? (double)pl->total_or * (double)pl->nsteps / s.elapsed_s : 0.0;
  s.full_dim = pl->full_dim;
  s.reduced_dim = pl->red_dim;
  s.reduced_nnz = pl->red_nnz;
  s.padded_dim = pl->np_max;
  s.block = pl->block;
  s.taylor_terms = pl->taylor_k;
  std::vector<int32_t> sq((size_t)pl->total_or);
  HIPCHK(hipMemcpy(sq.data(), pl->d_sq.p, sq.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  int mx = 0;
  int64_t sum = 0;
  for (int v : sq) {
    mx = std::max(mx, v);
    sum += v;
  }
  s.max_squarings = mx;
  s.ps_block = pl->ps_q;
  s.real_systems = pl->real_systems;
  s.spectral_systems = pl->spectral_systems;
  s.eig_max_sweeps = max_sweeps;
  s.krylov_systems = pl->krylov_systems;
  s.step_products = step_products;
  s.plan_ms = pl->plan_ms;
  for (int k = 0; k < 4; k++) s.stage_ms[k] = ms_stage[k];
  s.sum_squarings_lo = (int32_t)(sum & 0x7fffffff);
  s.eig_fallbacks = eig_failures;
  if (pl->mas_tiled) {
    std::vector<int32_t> np_or((size_t)pl->total_or);
    HIPCHK(hipMemcpy(np_or.data(), pl->d_nprod.p, np_or.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (int32_t v : np_or) tot += v;
    s.step_products = tot;
    s.mas_tiles = pl->mas_ntile;
    s.mas_pairs = pl->mas_npair;
    s.mas_detected = pl->mas_ldet;
  }
  pl->last = s;
  if (stats) *stats = s;
  pl->executed = true;  // sp_plan_result refuses the results of a failed pass
  return SP_OK;
}

// Device address of the static scratch, looked up once per device (the device
// must be current).
std::mutex g_scratch_mu;
void* g_scratch_dev[256] = {};

void* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= 256) throw SpError(SP_EINVAL, "device index out of range");
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = ptr;
  }
  return g_scratch_dev[device];
}

// Bump allocation inside the static scratch (256-byte aligned pieces).
struct Arena {
  unsigned char* base;
  size_t cap;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    const size_t off = (used + 255) & ~(size_t)255;
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (off > cap || bytes > cap - off)
      throw SpError(SP_ERANGE, "the problem does not fit the " + std::to_string(kScratchBytes >> 20) +
                                   " MiB static scratch of sp_powder_fid_noalloc");
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
};

template <typename T>
void put(T* dst, const std::vector<T>& v, hipStream_t s) {
  if (!v.empty()) HIPCHK(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// sp_powder_fid on the spectral path with every device buffer carved from the
// static scratch, orientations in blocks of a multiple of 64 (so the chunk
// partials add up in the same order as k_system_sum's).
int noalloc_fid(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* al, const double* be, const double* ga,
                const double* w, int64_t n_or, const double* rho0, const double* coil, double dt, int64_t nsteps,
                const sp_tols* tols, int mode, double* fid_out, double* per_out, sp_stats* stats) {
  if (!fid_out) throw SpError(SP_EINVAL, "fid_out must not be NULL");
  if (!ctx->scratch) throw SpError(SP_EDEVICE, "the static scratch is not available: " + ctx->scratch_err);
  const Tols tl = read_tols(tols);
  if (mode != SP_MODE_FP64)
    throw SpError(SP_EINVAL, "sp_powder_fid_noalloc runs the FP64 spectral path only (mode must be SP_MODE_FP64)");
  if (!(std::isfinite(dt)) || dt <= 0) throw SpError(SP_EINVAL, "dt must be positive and finite");
  if (nsteps < 0 || nsteps > (int64_t)1 << 30) throw SpError(SP_EINVAL, "invalid nsteps");
  SysPlan sp = plan_system(I, Q, al, be, ga, w, n_or, rho0, coil);
  if (!sp.real || sp.np > spk::kTriNP || (double)sp.n >= tl.krylov_tol)
    throw SpError(SP_EINVAL,
                  "sp_powder_fid_noalloc runs the spectral path only: the generator must be real at every "
                  "orientation and its coil-connected dimension at most " + std::to_string(spk::kTriNP) +
                      " and below krylov_tol");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t npts = nsteps + 1;
  Arena a{static_cast<unsigned char*>(ctx->scratch), kScratchBytes};
  spk::SysDesc d{};
  d.rlx_off = -1;
  d.rdiag_off = -1;
  d.n = sp.n;
  d.np = sp.np;
  d.nnz = sp.nnz;
  d.nq = (int32_t)sp.slots.size();
  d.n_or = n_or;
  d.real = 1;
  for (int k = 0; k < spk::kRowBlocks; k++)
    d.rb[k] = (int32_t)(std::lower_bound(sp.row.begin(), sp.row.end(), 16 * k) - sp.row.begin());
  std::vector<spk::SysDesc> dv(1, d);
  auto* d_sys = a.take<spk::SysDesc>(1);
  auto* d_row = a.take<int32_t>(sp.row.size());
  auto* d_col = a.take<int32_t>(sp.col.size());
  auto* d_tpos = a.take<int32_t>(sp.tpos.size());
  auto* d_vals = a.take<double>(sp.vals.size());
  auto* d_slots = a.take<int32_t>(sp.slots.size());
  auto* d_rho = a.take<double>(sp.rho.size());
  auto* d_coil = a.take<double>(sp.coil.size());
  auto* d_al = a.take<double>((size_t)n_or);
  auto* d_be = a.take<double>((size_t)n_or);
  auto* d_ga = a.take<double>((size_t)n_or);
  auto* d_w = a.take<double>((size_t)n_or);
  auto* d_fid = a.take<double>(2 * (size_t)npts);
  auto* d_queue = a.take<unsigned int>(spk::kQueues);
  // block of C orientations: items, chunk table, spectra, status, FIDs, partials
  const size_t per64 = 64 * (spk::kSpecStride * sizeof(double) + sizeof(int32_t) + sizeof(spk::Item) +
                             (size_t)npts * 2 * sizeof(double)) +
                       (size_t)npts * 2 * sizeof(double) + sizeof(spk::SumChunk);
  const size_t left = a.used + 8 * 256 < kScratchBytes ? kScratchBytes - a.used - 8 * 256 : 0;
  int64_t C = 64 * (int64_t)(left / per64);
  if (C < 64)
    throw SpError(SP_ERANGE, "nsteps is too large for the " + std::to_string(kScratchBytes >> 20) +
                                 " MiB static scratch of sp_powder_fid_noalloc (" + std::to_string(per64) +
                                 " bytes per 64 orientations)");
  C = std::min<int64_t>(C, (n_or + 63) / 64 * 64);
  auto* d_items = a.take<spk::Item>((size_t)C);
  auto* d_chunks = a.take<spk::SumChunk>((size_t)C / 64);
  auto* d_spec = a.take<double>((size_t)C * spk::kSpecStride);
  auto* d_status = a.take<int32_t>((size_t)C);
  auto* d_per = a.take<double>((size_t)C * npts * 2);
  auto* d_partial = a.take<double>((size_t)C / 64 * npts * 2);
  std::vector<double> alv(al, al + n_or), bev(be, be + n_or), gav(ga, ga + n_or), wv(w, w + n_or);
  put(d_sys, dv, st);
  put(d_row, sp.row, st);
  put(d_col, sp.col, st);
  put(d_tpos, sp.tpos, st);
  put(d_vals, sp.vals, st);
  put(d_slots, sp.slots, st);
  put(d_rho, sp.rho, st);
  put(d_coil, sp.coil, st);
  put(d_al, alv, st);
  put(d_be, bev, st);
  put(d_ga, gav, st);
  put(d_w, wv, st);
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventRecord(ctx->ev0, st));
  const unsigned nbt = (unsigned)((npts + 255) / 256);
  hipLaunchKernelGGL(spk::k_fill_zero, dim3((unsigned)((2 * npts + 255) / 256)), dim3(256), 0, st, d_fid, 2 * npts);
  HIPCHK(hipGetLastError());
  std::vector<spk::Item> items;
  std::vector<spk::SumChunk> chunks;
  std::vector<int32_t> status;
  int32_t max_sweeps = 0;
  for (int64_t b0 = 0; b0 < n_or; b0 += C) {
    const int64_t nb = std::min<int64_t>(C, n_or - b0);
    items.clear();
    chunks.clear();
    for (int64_t k = 0; k < nb; k++) items.push_back(spk::Item{0, 0, k});
    for (int64_t k = 0; k < nb; k += 64) chunks.push_back(spk::SumChunk{k, (int32_t)std::min<int64_t>(64, nb - k), 0});
    put(d_items, items, st);
    put(d_chunks, chunks, st);
    hipLaunchKernelGGL(spk::k_reset, dim3(1), dim3(64), 0, st, d_queue);
    HIPCHK(hipGetLastError());
    spk::SpecParams sp2{};
    sp2.sys = d_sys;
    sp2.items = d_items;
    sp2.n_items = nb;
    sp2.queue = d_queue + 1;
    sp2.prow = d_row;
    sp2.pcol = d_col;
    sp2.ptpos = d_tpos;
    sp2.vals = d_vals;
    sp2.slots = d_slots;
    sp2.rho0 = d_rho;
    sp2.coil = d_coil;
    sp2.alphas = d_al + b0;
    sp2.betas = d_be + b0;
    sp2.gammas = d_ga + b0;
    sp2.dt = dt;
    sp2.nsteps = nsteps;
    sp2.spec = d_spec;
    sp2.spec_nm = spk::kTriNP;
    sp2.status = d_status;
    sp2.per_or = d_per;
    const int g_tri = (int)std::max<int64_t>(1, std::min<int64_t>(2 * (int64_t)ctx->n_cu, nb));
    const int g_syn = (int)std::max<int64_t>(1, std::min<int64_t>(3 * (int64_t)ctx->n_cu, nb));
    hipLaunchKernelGGL(spk::k_tridiag, dim3(g_tri), dim3(spk::kTriThreads), 0, st, sp2);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(spk::k_dceig<spk::kTriNP>, dim3((unsigned)nb), dim3(spk::DcSharedT<spk::kTriNP>::kThr), 0, st,
                       sp2);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(spk::k_synth<spk::kTriNP>, dim3(g_syn), dim3(spk::kThreads), 0, st, sp2);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(spk::k_chunk_sum, dim3(nbt, (unsigned)chunks.size()), dim3(256), 0, st, d_per, d_w + b0,
                       d_chunks, npts, d_partial);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(spk::k_accumulate, dim3(nbt), dim3(256), 0, st, d_partial, (int32_t)chunks.size(), npts,
                       d_fid);
    HIPCHK(hipGetLastError());
    status.resize((size_t)nb);
    HIPCHK(hipMemcpyAsync(status.data(), d_status, (size_t)nb * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (per_out)
      HIPCHK(hipMemcpyAsync(per_out + 2 * (size_t)b0 * npts, d_per, (size_t)nb * npts * 2 * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));  // the next block reuses the block buffers
    for (int32_t v : status) {
      if (v < 0)
        throw SpError(SP_ERANGE, "the tridiagonal eigensolver did not converge for an orientation");
      max_sweeps = std::max(max_sweeps, v);
    }
  }
  HIPCHK(hipEventRecord(ctx->ev2, st));
  HIPCHK(hipMemcpyAsync(fid_out, d_fid, (size_t)npts * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stats) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev2);
    sp_stats s{};
    s.elapsed_s = ms * 1e-3;
    s.kernel_ms = ms;
    s.orient_steps_per_s = s.elapsed_s > 0 ? (double)n_or * (double)nsteps / s.elapsed_s : 0.0;
    s.full_dim = sp.full_dim;
    s.reduced_dim = sp.n;
    s.reduced_nnz = sp.nnz;
    s.padded_dim = sp.np;
    s.real_systems = 1;
    s.spectral_systems = 1;
    s.eig_max_sweeps = max_sweeps;
    *stats = s;
  }
  return SP_OK;
}

// ------------------------------------------------------------- autotuning
std::mutex g_tune_mu;
std::map<std::string, sp_tuning> g_tune_cache;

std::string tune_key(const sp_plan* pl) {
  char b[256];
  snprintf(b, sizeof(b), "m%d s%lld p%lld/%lld k%lld/%lld t%lld np%d/%d o%lld", pl->mode, (long long)pl->nsteps,
           (long long)pl->n_prop, (long long)pl->n_prop_complex, (long long)pl->n_spec, (long long)pl->n_kry,
           (long long)pl->n_systems, pl->np_prop, pl->np_max, (long long)pl->total_or);
  return b;
}

sp_tuning get_tuning(const sp_plan* pl) {
  sp_tuning t{};
  t.log2_block = pl->log2b;
  t.prop_wg_per_cu = pl->per_cu;
  t.tri_wg_per_cu = pl->tri_per_cu;
  t.syn_wg_per_cu = pl->syn_per_cu;
  t.step_wg_per_cu = pl->step_per_cu;
  return t;
}

void apply_tuning(sp_plan* pl, const sp_tuning& t) {
  sp_ctx* ctx = pl->ctx;
  if (pl->n_prop > 0 && (t.log2_block != pl->log2b || t.prop_wg_per_cu != pl->per_cu))
    configure_prop(ctx, pl, t.log2_block, t.prop_wg_per_cu);
  if (pl->n_spec > 0) configure_spec(ctx, pl, t.tri_wg_per_cu, t.syn_wg_per_cu);
  if (pl->step_lds[0] > 0 && t.step_wg_per_cu != pl->step_per_cu) configure_step(ctx, pl, t.step_wg_per_cu);
  if (pl->gexec) {  // the recorded pass holds the old grids and workspaces
    (void)hipGraphExecDestroy(pl->gexec);
    pl->gexec = nullptr;
  }
  pl->executed = false;
}

// Mean kernel time of `passes` eager passes after one warm-up.
double time_passes(sp_plan* pl, int passes) {
  const bool graph = pl->graph;
  pl->graph = false;
  sp_stats st{};
  double tot = 0.0;
  try {
    run(pl, &st);
    for (int k = 0; k < passes; k++) {
      run(pl, &st);
      tot += st.kernel_ms;
    }
  } catch (...) {
    pl->graph = graph;
    throw;
  }
  pl->graph = graph;
  return tot / passes;
}

int autotune(sp_plan* pl, int passes, sp_tuning* out) {
  if (pl->mas) throw SpError(SP_EINVAL, "MAS plans have no tunable launch configuration");
  HIPCHK(hipSetDevice(pl->ctx->device));
  const std::string key = tune_key(pl);
  if (passes >= 0) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      apply_tuning(pl, it->second);
      if (out) *out = it->second, out->candidates = 0;
      return SP_OK;
    }
  }
  const int np = std::max(1, passes < 0 ? -passes : passes);
  sp_tuning best = get_tuning(pl);
  double best_ms = time_passes(pl, np);
  const double default_ms = best_ms;
  int cand = 1;
  auto trial = [&](sp_tuning t) {
    apply_tuning(pl, t);
    const double ms = time_passes(pl, np);
    cand++;
    if (ms < best_ms) {
      best_ms = ms;
      best = get_tuning(pl);
    }
  };
  const int64_t npts = pl->nsteps + 1;
  if (pl->n_prop > 0) {
    const int l0 = best.log2_block;
    for (int l = std::max(0, l0 - 2); l <= std::min(10, l0 + 2); l++)
      if (l != l0 && ((int64_t)1 << l) <= std::max<int64_t>(1, npts)) {
        sp_tuning t = best;
        t.log2_block = l;
        trial(t);
      }
    const int p0 = best.prop_wg_per_cu;
    for (int pc = 1; pc <= 3; pc++)
      if (pc != p0) {
        sp_tuning t = best;
        t.prop_wg_per_cu = pc;
        trial(t);
      }
  }
  if (pl->n_spec > 0) {
    for (int tc = 1; tc <= 2; tc++)
      if (tc != best.tri_wg_per_cu) {
        sp_tuning t = best;
        t.tri_wg_per_cu = tc;
        trial(t);
      }
    for (int sc = 1; sc <= 4; sc++)
      if (sc != best.syn_wg_per_cu) {
        sp_tuning t = best;
        t.syn_wg_per_cu = sc;
        trial(t);
      }
  }
  if (pl->n_kry > 0)
    for (int kc = 1; kc <= 4; kc++)
      if (kc != best.step_wg_per_cu) {
        sp_tuning t = best;
        t.step_wg_per_cu = kc;
        trial(t);
      }
  apply_tuning(pl, best);
  best = get_tuning(pl);
  best.candidates = cand;
  best.default_ms = default_ms;
  best.best_ms = best_ms;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  if (out) *out = best;
  return SP_OK;
}

// A cached tuning for a newly built plan of the same shape, if any.
void apply_cached_tuning(sp_plan* pl) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_cache.find(tune_key(pl));
  if (it != g_tune_cache.end()) apply_tuning(pl, it->second);
}

}  // namespace

extern "C" {

sp_ctx* sp_create(int device) {
  try {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) throw SpError(SP_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw SpError(SP_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<sp_ctx> c(new sp_ctx);
    c->device = device;
    try {
      c->scratch = scratch_address(device);  // once per device, with the device current
    } catch (SpError& e) {
      c->scratch_err = e.what();  // only sp_powder_fid_noalloc needs it
    }
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    HIPCHK(hipEventCreate(&c->ev2));
    for (auto& e : c->evs) HIPCHK(hipEventCreate(&e));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    return c.release();
  } catch (std::exception& ex) {
    g_create_err = ex.what();
    return nullptr;
  }
}

void sp_destroy(sp_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  for (auto& e : ctx->evs)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int sp_generator_is_real(const sp_csc* I, const sp_csc* Q, int32_t* is_real) {
  try {
    if (!is_real) throw SpError(SP_EINVAL, "is_real must not be NULL");
    *is_real = generator_is_real_full(I, Q) ? 1 : 0;
    return SP_OK;
  } catch (SpError& e) {
    g_create_err = e.what();
    return e.code;
  } catch (std::bad_alloc&) {
    g_create_err = "host allocation failed";
    return SP_ENOMEM;
  } catch (std::exception& e) {
    g_create_err = e.what();
    return SP_EINVAL;
  }
}

struct sp_ham {
  std::vector<spb::Csc> m;
};

sp_ham* sp_ham_build(int32_t nspins, const int8_t* basis, int64_t dim, const sp_ham_term* terms, int32_t nterms,
                     double tol) {
  try {
    if (!basis || dim < 1 || nterms < 0 || (nterms > 0 && !terms)) throw SpError(SP_EINVAL, "empty basis or NULL terms");
    if (!(tol > 0) || !std::isfinite(tol)) throw SpError(SP_EINVAL, "tol must be positive and finite");
    std::vector<int8_t> digits(basis, basis + (size_t)dim * nspins);
    std::vector<spb::Term> tv((size_t)nterms);
    for (int32_t i = 0; i < nterms; i++) {
      const sp_ham_term& a = terms[i];
      spb::Term& t = tv[(size_t)i];
      t.nspin = a.nspin;
      for (int k = 0; k < 2; k++) {
        t.spin[k] = a.spin[k];
        t.ist[k] = a.ist[k];
      }
      t.coef = a.coef;
      t.iso = spb::cd(a.iso[0], a.iso[1]);
      for (int k = 0; k < 8; k++) {
        t.istc[k] = spb::cd(a.ist_coef[2 * k], a.ist_coef[2 * k + 1]);
        t.irr[k] = spb::cd(a.irr[2 * k], a.irr[2 * k + 1]);
      }
    }
    std::unique_ptr<sp_ham> h(new sp_ham);
    h->m = spb::build_hamiltonian(nspins, digits, dim, tv, tol);
    return h.release();
  } catch (SpError& e) {
    g_create_err = e.what();
  } catch (std::invalid_argument& e) {
    g_create_err = e.what();
  } catch (std::bad_alloc&) {
    g_create_err = "host allocation failed";
  } catch (std::exception& e) {
    g_create_err = e.what();
  }
  return nullptr;
}

int sp_ham_matrix(const sp_ham* h, int32_t which, sp_csc* out) {
  if (!h || !out || which < 0 || which >= (int32_t)h->m.size()) return SP_EINVAL;
  const spb::Csc& c = h->m[(size_t)which];
  out->n = c.n;
  out->jc = c.jc.data();
  out->ir = c.ir.data();
  out->pr = c.pr.data();
  out->is_complex = 1;
  return SP_OK;
}

void sp_ham_free(sp_ham* h) { delete h; }

const char* sp_last_error(const sp_ctx* ctx) {
  if (!ctx) return g_create_err.c_str();
  return ctx->err.c_str();
}

sp_plan* sp_plan_powder(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas, const double* betas,
                        const double* gammas, const double* weights, int64_t n_or, const double* rho0,
                        const double* coil, double dt, int64_t nsteps, const sp_tols* tols, int mode) {
  if (!ctx) return nullptr;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    const Tols tl = read_tols(tols);
    std::vector<SysPlan> sps;
    sps.push_back(plan_system(I, Q, alphas, betas, gammas, weights, n_or, rho0, coil));
    sp_plan* pl = build_plan(ctx, sps, dt, nsteps, mode, tl);
    pl->red_states = sps[0].states;
    pl->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return pl;
  } catch (SpError& e) {
    fail(ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    fail(ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    fail(ctx, SP_EDEVICE, e.what());
  }
  return nullptr;
}

sp_plan* sp_plan_powder_relax(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                              const double* betas, const double* gammas, const double* weights, int64_t n_or,
                              const double* rho0, const double* coil, double dt, int64_t nsteps, const sp_tols* tols,
                              int mode, const sp_relax* rlx) {
  if (!ctx) return nullptr;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    const Tols tl = read_tols(tols);
    std::vector<SysPlan> sps;
    sps.push_back(plan_system(I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, rlx));
    sp_plan* pl = build_plan(ctx, sps, dt, nsteps, mode, tl);
    pl->red_states = sps[0].states;
    pl->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return pl;
  } catch (SpError& e) {
    fail(ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    fail(ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    fail(ctx, SP_EDEVICE, e.what());
  }
  return nullptr;
}

sp_plan* sp_plan_powder_batch(sp_ctx* ctx, const sp_system* systems, int64_t n_systems, double dt, int64_t nsteps,
                              const sp_tols* tols, int mode) {
  if (!ctx) return nullptr;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    const Tols tl = read_tols(tols);
    if (!systems || n_systems <= 0) throw SpError(SP_EINVAL, "the batch must hold at least one system");
    std::vector<SysPlan> sps;
    for (int64_t s = 0; s < n_systems; s++) {
      const sp_system& y = systems[s];
      sps.push_back(plan_system(y.I, y.Q, y.alphas, y.betas, y.gammas, y.weights, y.n_or, y.rho0, y.coil));
    }
    sp_plan* pl = build_plan(ctx, sps, dt, nsteps, mode, tl);
    for (int64_t s = 0; s < n_systems; s++) pl->fid_targets.push_back(systems[s].fid_out);
    pl->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return pl;
  } catch (SpError& e) {
    fail(ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    fail(ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    fail(ctx, SP_EDEVICE, e.what());
  }
  return nullptr;
}

int sp_execute(sp_plan* plan, sp_stats* stats) {
  if (!plan) return SP_EINVAL;
  try {
    return run(plan, stats);
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_result(sp_plan* plan, double* fid_out, double* per_orient_out) {
  if (!plan) return SP_EINVAL;
  try {
    if (!plan->executed) throw SpError(SP_EINVAL, "sp_execute has not run on this plan");
    HIPCHK(hipSetDevice(plan->ctx->device));
    const int64_t npts = plan->nsteps + 1;
    if (fid_out)
      HIPCHK(hipMemcpy(fid_out, plan->d_fid.p, (size_t)npts * 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (per_orient_out)
      HIPCHK(hipMemcpy(per_orient_out, plan->d_per.p, (size_t)plan->n_or[0] * npts * 2 * sizeof(double),
                       hipMemcpyDeviceToHost));
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_result_batch(sp_plan* plan) {
  if (!plan) return SP_EINVAL;
  try {
    if (!plan->executed) throw SpError(SP_EINVAL, "sp_execute has not run on this plan");
    if ((int64_t)plan->fid_targets.size() != plan->n_systems) throw SpError(SP_EINVAL, "not a batch plan");
    HIPCHK(hipSetDevice(plan->ctx->device));
    const int64_t npts = plan->nsteps + 1;
    std::vector<double> all((size_t)plan->n_systems * npts * 2);
    HIPCHK(hipMemcpy(all.data(), plan->d_fid.p, all.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < plan->n_systems; s++)
      if (plan->fid_targets[s])
        std::memcpy(plan->fid_targets[s], all.data() + 2 * (size_t)s * npts, (size_t)npts * 2 * sizeof(double));
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_set_dead_time(sp_plan* plan, double dead_time) {
  if (!plan) return SP_EINVAL;
  try {
    if (plan->mas) throw SpError(SP_EINVAL, "a dead time is not available on MAS plans");
    if (!std::isfinite(dead_time) || dead_time < 0) throw SpError(SP_EINVAL, "dead_time must be finite and >= 0");
    HIPCHK(hipSetDevice(plan->ctx->device));
    if (dead_time > 0 && plan->init_states)
      throw SpError(SP_EINVAL, "a dead time and per-orientation initial states cannot be combined");
    if (dead_time > 0) {
      if (plan->step_lds[0] == 0) setup_step(plan->ctx, plan);
      if (!plan->d_rho_or.p)
        plan->d_rho_or.alloc((size_t)plan->total_or * plan->rho_or_stride * 2 * sizeof(double));
      if (!plan->d_nsub_fs.p) plan->d_nsub_fs.alloc((size_t)plan->total_or * sizeof(int32_t));
    }
    if (dead_time != plan->dead_time && plan->gexec) {  // the recorded pass changes
      (void)hipGraphExecDestroy(plan->gexec);
      plan->gexec = nullptr;
    }
    plan->dead_time = dead_time;
    plan->executed = false;
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_set_initial_states(sp_plan* plan, const double* rho_or) {
  if (!plan) return SP_EINVAL;
  try {
    if (plan->mas) throw SpError(SP_EINVAL, "initial states are not available on MAS plans");
    if (plan->n_systems != 1 || plan->red_states.empty())
      throw SpError(SP_EINVAL, "per-orientation initial states need a single-system plan");
    HIPCHK(hipSetDevice(plan->ctx->device));
    const bool was = plan->init_states;
    if (!rho_or) {
      plan->init_states = false;
    } else {
      if (plan->dead_time > 0)
        throw SpError(SP_EINVAL, "a dead time and per-orientation initial states cannot be combined");
      const int64_t dim = plan->full_dim, n = (int64_t)plan->red_states.size(), s = plan->rho_or_stride;
      std::vector<double> red((size_t)plan->total_or * s * 2, 0.0);
      for (int64_t o = 0; o < plan->total_or; o++)
        for (int64_t k = 0; k < n; k++) {
          const double re = rho_or[2 * (o * dim + plan->red_states[k])];
          const double im = rho_or[2 * (o * dim + plan->red_states[k]) + 1];
          if (!std::isfinite(re) || !std::isfinite(im)) throw SpError(SP_EINVAL, "non-finite initial state");
          red[2 * (o * s + k)] = re;
          red[2 * (o * s + k) + 1] = im;
        }
      if (!plan->d_rho_or.p) plan->d_rho_or.alloc(red.size() * sizeof(double));
      HIPCHK(hipMemcpyAsync(plan->d_rho_or.p, red.data(), red.size() * sizeof(double), hipMemcpyHostToDevice,
                            plan->ctx->stream));
      HIPCHK(hipStreamSynchronize(plan->ctx->stream));
      plan->init_states = true;
    }
    if (was != plan->init_states && plan->gexec) {  // the recorded pass reads other states
      (void)hipGraphExecDestroy(plan->gexec);
      plan->gexec = nullptr;
    }
    plan->executed = false;
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_propagate(sp_plan* plan, double dt, int64_t nsteps, double* rho_out) {
  if (!plan) return SP_EINVAL;
  try {
    if (plan->mas) throw SpError(SP_EINVAL, "propagation of states is not available on MAS plans");
    if (plan->n_systems != 1 || plan->red_states.empty())
      throw SpError(SP_EINVAL, "propagation of states needs a single-system plan");
    if (plan->dead_time > 0) throw SpError(SP_EINVAL, "propagation of states does not apply the dead time");
    if (!std::isfinite(dt) || dt <= 0) throw SpError(SP_EINVAL, "dt must be positive and finite");
    if (nsteps < 1 || nsteps > ((int64_t)1 << 30)) throw SpError(SP_EINVAL, "nsteps must be between 1 and 2^30");
    if (!rho_out) throw SpError(SP_EINVAL, "rho_out is NULL");
    sp_ctx* ctx = plan->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    if (plan->step_lds[0] == 0) setup_step(ctx, plan);
    const int64_t s = plan->rho_or_stride;
    if (!plan->d_prop_out.p) plan->d_prop_out.alloc((size_t)plan->total_or * s * 2 * sizeof(double));
    if (!plan->d_nsub_fs.p) plan->d_nsub_fs.alloc((size_t)plan->total_or * sizeof(int32_t));
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(spk::k_reset, dim3(1), dim3(64), 0, st, plan->d_queue.as<unsigned int>());
    HIPCHK(hipGetLastError());
    launch_step_items(plan, false, plan->d_items.as<spk::Item>(), plan->n_items, 6, true, dt, nsteps,
                      plan->d_prop_out.as<double>(), plan->init_states ? plan->d_rho_or.as<double>() : nullptr);
    std::vector<double> red((size_t)plan->total_or * s * 2);
    std::vector<int32_t> ns((size_t)plan->total_or);
    HIPCHK(hipMemcpyAsync(red.data(), plan->d_prop_out.p, red.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ns.data(), plan->d_nsub_fs.p, ns.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int32_t v : ns)
      if (v > spk::kStepMaxSub)
        throw SpError(SP_ERANGE, "either dt is too long, or L is too big: |L*dt|>1e4, check both.");
    const int64_t dim = plan->full_dim, n = (int64_t)plan->red_states.size();
    std::fill(rho_out, rho_out + 2 * plan->total_or * dim, 0.0);
    for (int64_t o = 0; o < plan->total_or; o++)
      for (int64_t k = 0; k < n; k++) {
        rho_out[2 * (o * dim + plan->red_states[k])] = red[2 * (o * s + k)];
        rho_out[2 * (o * dim + plan->red_states[k]) + 1] = red[2 * (o * s + k) + 1];
      }
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    return fail(plan->ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_autotune(sp_plan* plan, int32_t passes, sp_tuning* chosen) {
  if (!plan) return SP_EINVAL;
  try {
    return autotune(plan, passes, chosen);
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

int sp_plan_get_tuning(const sp_plan* plan, sp_tuning* out) {
  if (!plan || !out) return SP_EINVAL;
  *out = get_tuning(plan);
  return SP_OK;
}

int sp_plan_set_tuning(sp_plan* plan, const sp_tuning* in) {
  if (!plan || !in) return SP_EINVAL;
  try {
    if (plan->mas) throw SpError(SP_EINVAL, "MAS plans have no tunable launch configuration");
    HIPCHK(hipSetDevice(plan->ctx->device));
    apply_tuning(plan, *in);
    return SP_OK;
  } catch (SpError& e) {
    return fail(plan->ctx, e.code, e.what());
  } catch (std::exception& e) {
    return fail(plan->ctx, SP_EDEVICE, e.what());
  }
}

void sp_plan_destroy(sp_plan* plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  if (plan->gexec) (void)hipGraphExecDestroy(plan->gexec);
  delete plan;
}

sp_plan* sp_plan_mas(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas, const double* betas,
                     const double* gammas, const double* weights, int64_t n_or, const double* rho0, const double* coil,
                     const sp_rotor* rotor, int64_t nsteps, int mode) {
  if (!ctx) return nullptr;
  try {
    return build_mas_plan(ctx, I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, rotor, nsteps, mode);
  } catch (SpError& e) {
    fail(ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    fail(ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    fail(ctx, SP_EDEVICE, e.what());
  }
  return nullptr;
}

int sp_mas_fid(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas, const double* betas,
               const double* gammas, const double* weights, int64_t n_or, const double* rho0, const double* coil,
               const sp_rotor* rotor, int64_t nsteps, int mode, double* fid_out, double* per_orient_out,
               sp_stats* stats) {
  if (!ctx) return SP_EINVAL;
  if (!fid_out) return fail(ctx, SP_EINVAL, "fid_out must not be NULL");
  ctx->code = 0;
  sp_plan* pl = sp_plan_mas(ctx, I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, rotor, nsteps, mode);
  if (!pl) return ctx->code ? ctx->code : SP_EDEVICE;
  int rc = sp_execute(pl, stats);
  if (rc == SP_OK) rc = sp_plan_result(pl, fid_out, per_orient_out);
  sp_plan_destroy(pl);
  return rc;
}

int sp_powder_fid(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas, const double* betas,
                  const double* gammas, const double* weights, int64_t n_or, const double* rho0, const double* coil,
                  double dt, int64_t nsteps, const sp_tols* tols, int mode, double* fid_out, double* per_orient_out,
                  sp_stats* stats) {
  if (!ctx) return SP_EINVAL;
  if (!fid_out) return fail(ctx, SP_EINVAL, "fid_out must not be NULL");
  ctx->code = 0;
  sp_plan* pl = sp_plan_powder(ctx, I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, dt, nsteps, tols, mode);
  if (!pl) return ctx->code ? ctx->code : SP_EDEVICE;
  int rc = sp_execute(pl, stats);
  if (rc == SP_OK) rc = sp_plan_result(pl, fid_out, per_orient_out);
  sp_plan_destroy(pl);
  return rc;
}

int sp_powder_fid_noalloc(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                          const double* betas, const double* gammas, const double* weights, int64_t n_or,
                          const double* rho0, const double* coil, double dt, int64_t nsteps, const sp_tols* tols,
                          int mode, double* fid_out, double* per_orient_out, sp_stats* stats) {
  if (!ctx) return SP_EINVAL;
  try {
    return noalloc_fid(ctx, I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, dt, nsteps, tols, mode, fid_out,
                       per_orient_out, stats);
  } catch (SpError& e) {
    return fail(ctx, e.code, e.what());
  } catch (std::bad_alloc&) {
    return fail(ctx, SP_ENOMEM, "host allocation failed");
  } catch (std::exception& e) {
    return fail(ctx, SP_EDEVICE, e.what());
  }
}

}  // extern "C"

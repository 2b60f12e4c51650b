// Time stepping by the reordered Taylor series: the step() / krylov() path.
//
// kernel/krylov.m:139-163 ('observable') propagates each subspace one time step
// at a time with step(L,rho,dt) (kernel/step.m:204-267, reordered_taylor
// :328-375) and detects coil'*rho after every step; evolution.m:574 takes it
// for subspaces of dimension >= krylov_tol or when sys.enable holds 'krylov'.
// experiments/acquire.m:65-67 runs the same step() once over the dead time.
//
// One 256-thread workgroup per (system, orientation) item, pulled from a
// persistent queue. Per item:
//   1. the rank-1/rank-2 Wigner coefficients of the orientation,
//   2. L = (x + x')/2, x = I + sum_q c_q Q_q on the reduced union pattern,
//      written in sliced-ELL layout: a slice of 64/G rows is read by one wave,
//      G lanes per row (G = 1, 2, 4 or 8 from the mean row length), entry k of
//      the slice's row r at slice_offset + 64(k/G) + G r + k%G, so that one
//      load instruction reads 64 consecutive entries; real values
//      when the generator is real at every orientation, complex otherwise;
//      staged in LDS when the plan's largest slice layout fits, else in the
//      workgroup's slot of a global workspace,
//   3. the largest row sum of |L| in FP64 (cheap_norm.m:45-50; equal to the
//      1-norm for a Hermitian L, and with relaxation a bound of the same use:
//      each substep's product has norm <= 2 in it) and the substep count
//      nsub = ceil(|L|_1 |dt| / 2) (step.m:222-227),
//   4. per time step: scaling = max|rho| (a workgroup max reduction),
//      rho /= scaling; nsub substeps of next = -(i/k)(dt/nsub) L next,
//      rho += next, k = 1, 2, ... while any |next| > eps (one
//      __syncthreads_or per term: the series' own convergence test);
//      rho *= scaling; then f(t) = coil'*rho (a workgroup sum).
// Row ownership follows the slices: slice s belongs to wave s % 4 and its row
// r' to lane G r' of that wave (the row's partial sums meet there by cross-lane
// shuffles), so every thread updates only the rows it owns and the state
// vector it reads in the product is double-buffered in LDS (one barrier per
// Taylor term).
#pragma once

#include "powder_kernels.hpp"

namespace spk {

constexpr int kStepThreads = 256;
constexpr int kSell = 64;           // entries per slice row of the layout (one wave)
constexpr int kStepMaxSub = 10000;  // substeps per time step (step.m:230-233)
constexpr int kStepMaxTerms = 256;  // Taylor terms per substep: |L dt/nsub|_1 <= 2 converges in ~30
constexpr int kStepMaxDim = 2304;                  // largest coil-connected dimension of the path
constexpr int kStepMaxSlices = kStepMaxDim / 8;     // slices hold at least 8 rows (at most 8 lanes per row)
// register-cached product: a wave holds RCS (1 or 2) slices of at most kRcEntries / RCS
// 64-entry columns each; its lanes keep their columns and values in registers for the item
constexpr int kRcEntries = 18;
__host__ __device__ constexpr int rc_width(int rcs) { return rcs > 0 ? kRcEntries / rcs : 1; }

// Per-system layout of the sliced-ELL generator.
struct StepDesc {
  int32_t nslices;    // ceil(n / (64 / G))
  int32_t sell_nnz;   // entries of the slice layout, padding included (64 * sum of slice widths)
  int64_t slice_off;  // offset into the slice-start array (nslices + 1 entries)
  int64_t sell_off;   // offset into the slice-layout column array (sell_nnz entries)
  int32_t lanes;      // G, lanes per row
  int32_t log2_lanes;
  int32_t wmax;       // widest slice, in 64-entry columns
  int32_t pad;
  int64_t perm_off;   // offset of the system's state order in sell_perm: layout row k is state sell_perm[k]
};

struct StepParams {
  const SysDesc* sys;
  const StepDesc* sdesc;
  const Item* items;
  int64_t n_items;
  unsigned int* queue;    // work-queue head, zeroed by k_reset
  const int32_t* ptpos;   // transposed pattern entry (union pattern, row-sorted)
  const double* vals;     // interleaved complex FP64 value planes: I, then each non-empty Q slot
  const int32_t* slots;   // Q slot ids
  const int32_t* sell_start;  // per system: start of each slice within the system's slice layout
  const int32_t* sell_col;    // column of every slice-layout entry (padding: the entry's own row)
  const int32_t* sell_pos;    // slice-layout position of every union-pattern entry
  const int32_t* sell_perm;   // per system: the reduced state of every layout row
  const double* rho0;     // interleaved complex, per system
  const double* rho_or;   // optional per-orientation initial states (dead time), stride rho_or_stride
  int64_t rho_or_stride;  // complex entries per orientation in rho_or / rho_out
  const double* coil;
  const double* alphas;
  const double* betas;
  const double* gammas;
  double dt;              // time step
  int64_t nsteps;         // steps; the observable has nsteps+1 points
  int32_t final_state;    // 0: write f(t) to per_or; 1: write rho after nsteps steps to rho_out
  int32_t vec_words;      // doubles of the four state vectors (8 * largest n, rounded up to 64)
  int32_t vec_global;     // 1: rho and the coil copy live in the workgroup's slot at gvec_off
  int64_t gvec_off;       //    (only the two term buffers in LDS: twice the workgroups per CU)
  double* work;           // per-workgroup slots: values not staged in LDS, global state vectors
  int64_t work_stride;    // doubles per workgroup slot
  double* per_or;         // per-orientation FIDs, interleaved complex, nsteps+1 per orientation
  double* rho_out;        // final states (final_state == 1)
  int32_t* nsub_out;      // per orientation: substeps per time step (diagnostics)
  int32_t* terms_out;     // per orientation: Taylor terms (matrix-vector products) over all steps
  const int32_t* prow;    // row of every union-pattern entry (reduced state)
  const double* rvals;    // relaxation R0 on the union pattern (SysDesc::rlx_off), interleaved complex
  const double* rdiag;    // relaxation diagonal planes g0..g6 (SysDesc::rdiag_off)
};

// Sum and max over the workgroup: DPP within each wave, then four partials in LDS.
__device__ __forceinline__ double step_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double step_wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// VL: the slice-layout values and columns are staged in LDS (else read from the
// workgroup's global slot and the system's column array). RCS > 0 (with VL, for systems of
// at most 4*RCS slices no wider than rc_width(RCS)): each lane copies its entries' columns
// and values into registers once per item, so a term's product issues all its state gathers
// at once instead of a column load before each gather.
template <bool REAL, bool VL, int RCS = 0>
__global__ void __launch_bounds__(kStepThreads, 2) k_step(StepParams p) {  // two waves per SIMD: <= 256 VGPRs
  constexpr bool RC = RCS > 0;
  constexpr int kRcSlices = RC ? RCS : 1, kRcWidth = rc_width(RCS);
  static_assert(!RC || VL, "the register-cached product reads the staged layout");
  extern __shared__ __align__(16) double lds[];  // every region starts 16-byte aligned (double2 reads)
  __shared__ double cre[34], cim[34];
  __shared__ double red[2][kStepThreads / 64];
  __shared__ double rk[kStepMaxTerms + 1];  // 1/k
  __shared__ int sany[2][kStepThreads / 64];  // per-wave "a term entry exceeds eps", by term parity
  __shared__ int sst[kStepMaxSlices + 1];     // slice starts of the current item
  __shared__ int sitem;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int kVW = REAL ? 1 : 2;  // doubles per generator value
  const int q4 = p.vec_words / 4;    // doubles of one state vector
  double* wslot = p.work + (size_t)blockIdx.x * p.work_stride;
  double* va = lds;                  // next term, read buffer (interleaved complex)
  double* vb = lds + q4;             // next term, write buffer
  // rho: between barriers each thread touches only its own rows, so it may live in the
  // workgroup's global slot as well as in LDS; the coil copy after it
  double* vr = p.vec_global ? wslot + p.gvec_off : lds + 2 * q4;
  double* vc = vr + q4;              // conj-free coil copy
  double* lval = lds + (p.vec_global ? 2 : 4) * q4;  // staged values (VL)
  const double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
  for (int k = tid; k <= kStepMaxTerms; k += kStepThreads) rk[k] = k > 0 ? 1.0 / (double)k : 0.0;

  for (;;) {
    if (tid == 0) sitem = (int)atomicAdd(p.queue, 1u);
    __syncthreads();
    const int64_t it = __builtin_amdgcn_readfirstlane(sitem);
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sd = p.sys[item.sys];
    const StepDesc st = p.sdesc[item.sys];
    const int n = __builtin_amdgcn_readfirstlane(sd.n);
    const int nslices = __builtin_amdgcn_readfirstlane(st.nslices);
    const int sell_nnz = __builtin_amdgcn_readfirstlane(st.sell_nnz);
    const int lg = __builtin_amdgcn_readfirstlane(st.log2_lanes);
    const int G = 1 << lg, H = kSell >> lg;  // lanes per row, rows per slice
    const int rsub = lane >> lg;             // the lane's row within a slice
    const bool owner = (lane & (G - 1)) == 0;
    const int32_t* sstart = p.sell_start + st.slice_off;
    const int32_t* gcol = p.sell_col + st.sell_off;
    const int32_t* spos = p.sell_pos + sd.pat_off;
    const int64_t o = item.orient;
    double* val = VL ? lval : wslot;
    int32_t* lcol = reinterpret_cast<int32_t*>(lval + (size_t)kVW * sell_nnz);
    const int32_t* col = VL ? lcol : gcol;

    // 1. Wigner coefficients; clear the value layout (padding stays zero); stage columns
    if (tid < 34) wigner_coeff_slot(p.alphas[o], p.betas[o], p.gammas[o], tid, cre, cim);
    for (int k = tid; k < kVW * sell_nnz; k += kStepThreads) val[k] = 0.0;
    if (VL)
      for (int k = tid; k < sell_nnz; k += kStepThreads) lcol[k] = gcol[k];
    for (int k = tid; k <= nslices; k += kStepThreads) sst[k] = sstart[k];
    __syncthreads();
    // 2. L = (x + x')/2 on the union pattern, scattered into the slice layout
    {
      const double* v0 = p.vals + 2 * sd.val_off;
      const int32_t* slots = p.slots + sd.slot_off;
      const int32_t* ptp = p.ptpos + sd.pat_off;
      const int nnz = sd.nnz;
      double rq[6] = {0, 0, 0, 0, 0, 0};  // q_k(Omega) of the relaxation rate tensors
      if (!REAL && sd.rdiag_off >= 0) relax_monomials(p.betas[o], p.gammas[o], rq);
      for (int e = tid; e < nnz; e += kStepThreads) {
        const int et = ptp[e];
        double xr = v0[2 * e], xi = v0[2 * e + 1];
        double tr = v0[2 * et], ti = v0[2 * et + 1];
        for (int q = 0; q < sd.nq; q++) {
          const double* qv = v0 + 2 * (size_t)(q + 1) * nnz;
          const double cr = cre[slots[q]], ci = cim[slots[q]];
          xr += cr * qv[2 * e] - ci * qv[2 * e + 1];
          xi += cr * qv[2 * e + 1] + ci * qv[2 * e];
          tr += cr * qv[2 * et] - ci * qv[2 * et + 1];
          ti += cr * qv[2 * et + 1] + ci * qv[2 * et];
        }
        const int pos = spos[e];
        if (REAL) {
          val[pos] = 0.5 * (xr + tr);
        } else {
          // L = H + iR (relaxation.m, powder.m:263): R0 as given (not symmetrised) and, on the
          // diagonal, the orientation's rates g0 + sum_k q_k g_k
          double lr = 0.5 * (xr + tr), li = 0.5 * (xi - ti);
          if (sd.rlx_off >= 0) {
            lr -= p.rvals[2 * (sd.rlx_off + e) + 1];
            li += p.rvals[2 * (sd.rlx_off + e)];
          }
          if (sd.rdiag_off >= 0 && et == e) {
            const int i = p.prow[sd.pat_off + e];
            const double* g = p.rdiag + sd.rdiag_off;
            double r = g[i];
#pragma unroll
            for (int k = 0; k < 6; k++) r = fma(rq[k], g[(size_t)(k + 1) * sd.n + i], r);
            li += r;
          }
          val[2 * pos] = lr;
          val[2 * pos + 1] = li;
        }
      }
    }
    // states: rho0 (per system, or per orientation after a dead-time step) and the coil
    {
      const double* r0 = p.rho_or ? p.rho_or + 2 * (size_t)o * p.rho_or_stride : p.rho0 + 2 * sd.vec_off;
      const double* c0 = p.coil + 2 * sd.vec_off;
      const int32_t* perm = p.sell_perm + st.perm_off;
      for (int i = tid; i < n; i += kStepThreads) {  // layout row i holds state perm[i]
        const int q = perm[i];
        vr[2 * i] = r0[2 * q];
        vr[2 * i + 1] = r0[2 * q + 1];
        vc[2 * i] = c0[2 * q];
        vc[2 * i + 1] = c0[2 * q + 1];
      }
    }
    __syncthreads();
    // the lane's entries in registers: slice wave + 4j, column k (padding: value 0, own row)
    const int nsw = (nslices - wave + 3) >> 2;  // slices of this wave (uniform)
    int rc_col[kRcSlices][kRcWidth];
    double rc_val[kRcSlices][kRcWidth][REAL ? 1 : 2];
    if constexpr (RC) {
#pragma unroll
      for (int j = 0; j < kRcSlices; j++) {
        const int s = wave + 4 * j;
        const int b = (j < nsw) ? sst[s] : 0, w = (j < nsw) ? (sst[s + 1] - sst[s]) >> 6 : 0;
#pragma unroll
        for (int k = 0; k < kRcWidth; k++) {
          const bool in = k < w;
          const int e = b + (k << 6) + lane;
          rc_col[j][k] = in ? col[e] : 0;
          rc_val[j][k][0] = in ? val[kVW * e] : 0.0;
          if (!REAL) rc_val[j][k][REAL ? 0 : 1] = in ? val[kVW * e + 1] : 0.0;
        }
      }
    }
    // 3. |L|_1 = max row sum of |L_ij| (Hermitian L), substep count
    int nsub;
    {
      double mx = 0.0;
      for (int s = wave; s < nslices; s += 4) {
        const int b = sst[s], w = (sst[s + 1] - b) >> 6;
        double acc = 0.0;
        for (int k = 0; k < w; k++) {
          const int e = b + (k << 6) + lane;
          acc += REAL ? fabs(val[e]) : hypot(val[2 * e], val[2 * e + 1]);
        }
        for (int off = G >> 1; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);  // the row's G lanes
        mx = fmax(mx, acc);
      }
      mx = step_wave_max(mx);
      if (lane == 0) red[0][wave] = mx;
      __syncthreads();
      const double nrm = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3])) * fabs(p.dt);
      // step.m:230-233 stops above 1e4 substeps; the host turns the recorded count into that error
      nsub = (nrm <= 2.0 * kStepMaxSub) ? (int)ceil(nrm / 2.0) : kStepMaxSub + 1;
      if (tid == 0 && p.nsub_out) p.nsub_out[o] = nsub;
      __syncthreads();
    }
    if (nsub > kStepMaxSub) continue;
    const int64_t npts = p.nsteps + 1;
    double* out = p.per_or + 2 * (size_t)o * npts;
    // detection f = coil'*rho = sum conj(c_i) rho_i over the owned rows, then the workgroup
    auto detect = [&](int64_t t) {
      double sr = 0.0, si = 0.0;
      for (int s = wave; s < nslices; s += 4) {
        const int i = s * H + rsub;
        if (owner && i < n) {
          const double cr = vc[2 * i], ci = vc[2 * i + 1], rr = vr[2 * i], ri = vr[2 * i + 1];
          sr += cr * rr + ci * ri;
          si += cr * ri - ci * rr;
        }
      }
      sr = step_wave_sum(sr);
      si = step_wave_sum(si);
      if (lane == 0) {
        red[0][wave] = sr;
        red[1][wave] = si;
      }
      __syncthreads();
      if (tid == 0) {
        out[2 * t] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        out[2 * t + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
      }
      __syncthreads();
    };
    if (!p.final_state) detect(0);
    int32_t terms = 0;  // uniform: counted by every thread
    // 4. time steps
    for (int64_t t = 1; t <= p.nsteps; t++) {
      // scaling = max|rho| over the state (step.m:212-222)
      double m = 0.0;
      for (int s = wave; s < nslices; s += 4) {
        const int i = s * H + rsub;
        if (owner && i < n) m = fmax(m, hypot(vr[2 * i], vr[2 * i + 1]));
      }
      m = step_wave_max(m);
      if (lane == 0) red[0][wave] = m;
      __syncthreads();
      const double scaling = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
      __syncthreads();
      if (scaling > 0.0 && nsub > 0) {
        for (int s = wave; s < nslices; s += 4) {
          const int i = s * H + rsub;
          if (owner && i < n) {
            vr[2 * i] /= scaling;
            vr[2 * i + 1] /= scaling;
          }
        }
        const double h = p.dt / (double)nsub;
        for (int sub = 0; sub < nsub; sub++) {
          // next = rho on the owned rows; padding rows of the slices stay zero in both buffers
          for (int s = wave; s < nslices; s += 4) {
            const int i = s * H + rsub;
            if (owner) {
              double a = 0.0, b = 0.0;
              if (i < n) {
                a = vr[2 * i];
                b = vr[2 * i + 1];
              }
              va[2 * i] = a;
              va[2 * i + 1] = b;
              vb[2 * i] = 0.0;
              vb[2 * i + 1] = 0.0;
            }
          }
          __syncthreads();
          double* x = va;
          double* y = vb;
          for (int k = 1;; k++) {
            // next = -(1i/k)*(dt/nsub)*(L*next): with c = -(dt/nsub)/k, next = 1i*c*(L*next)
            const double c = -h * rk[k];
            int any = 0;
            const double2* x2 = reinterpret_cast<const double2*>(x);
            if constexpr (RC) {
              // every gather of the term first, then the row sums of each slice
              double2 g[kRcSlices][kRcWidth];
#pragma unroll
              for (int j = 0; j < kRcSlices; j++)
#pragma unroll
                for (int kk = 0; kk < kRcWidth; kk++)
                  if (j < nsw) g[j][kk] = x2[rc_col[j][kk]];  // padding: column 0, value 0
#pragma unroll
              for (int j = 0; j < kRcSlices; j++) {
                if (j < nsw) {
                  double tr = 0.0, ti = 0.0, ur = 0.0, ui = 0.0;  // two chains of the row sum
#pragma unroll
                  for (int kk = 0; kk < kRcWidth; kk++) {
                    {
                      const double2 pv = g[j][kk];
                      double& ar = (kk & 1) ? ur : tr;
                      double& ai = (kk & 1) ? ui : ti;
                      if (REAL) {
                        const double l0 = rc_val[j][kk][0];
                        ar = fma(l0, pv.x, ar);
                        ai = fma(l0, pv.y, ai);
                      } else {
                        const double lr = rc_val[j][kk][0], li = rc_val[j][kk][REAL ? 0 : 1];
                        ar = fma(lr, pv.x, ar);
                        ar = fma(-li, pv.y, ar);
                        ai = fma(lr, pv.y, ai);
                        ai = fma(li, pv.x, ai);
                      }
                    }
                  }
                  tr += ur;
                  ti += ui;
                  for (int off = G >> 1; off >= 1; off >>= 1) {  // the row's G lanes
                    tr += __shfl_xor(tr, off, 64);
                    ti += __shfl_xor(ti, off, 64);
                  }
                  const int i = (wave + 4 * j) * H + rsub;
                  const double nr = -c * ti, ni = c * tr;
                  if (owner) {
                    y[2 * i] = nr;
                    y[2 * i + 1] = ni;
                    if (i < n) {
                      vr[2 * i] += nr;
                      vr[2 * i + 1] += ni;
                      any |= (nr * nr + ni * ni > eps2) ? 1 : 0;
                    }
                  }
                }
              }
            } else
            for (int s = wave; s < nslices; s += 4) {
              const int b = sst[s], w = (sst[s + 1] - b) >> 6;
              double tr = 0.0, ti = 0.0, ur = 0.0, ui = 0.0;  // two chains of the row sum
              int kk = 0;
              for (; kk + 1 < w; kk += 2) {
                const int e0 = b + (kk << 6) + lane, e1 = e0 + 64;
                const double2 p0 = x2[col[e0]], p1 = x2[col[e1]];
                if (REAL) {
                  const double l0 = val[e0], l1 = val[e1];
                  tr = fma(l0, p0.x, tr);
                  ti = fma(l0, p0.y, ti);
                  ur = fma(l1, p1.x, ur);
                  ui = fma(l1, p1.y, ui);
                } else {
                  const double2 l0 = reinterpret_cast<const double2*>(val)[e0];
                  const double2 l1 = reinterpret_cast<const double2*>(val)[e1];
                  tr = fma(l0.x, p0.x, tr);
                  tr = fma(-l0.y, p0.y, tr);
                  ti = fma(l0.x, p0.y, ti);
                  ti = fma(l0.y, p0.x, ti);
                  ur = fma(l1.x, p1.x, ur);
                  ur = fma(-l1.y, p1.y, ur);
                  ui = fma(l1.x, p1.y, ui);
                  ui = fma(l1.y, p1.x, ui);
                }
              }
              if (kk < w) {
                const int e0 = b + (kk << 6) + lane;
                const double2 p0 = x2[col[e0]];
                if (REAL) {
                  const double l0 = val[e0];
                  tr = fma(l0, p0.x, tr);
                  ti = fma(l0, p0.y, ti);
                } else {
                  const double2 l0 = reinterpret_cast<const double2*>(val)[e0];
                  tr = fma(l0.x, p0.x, tr);
                  tr = fma(-l0.y, p0.y, tr);
                  ti = fma(l0.x, p0.y, ti);
                  ti = fma(l0.y, p0.x, ti);
                }
              }
              tr += ur;
              ti += ui;
              for (int off = G >> 1; off >= 1; off >>= 1) {  // the row's G lanes
                tr += __shfl_xor(tr, off, 64);
                ti += __shfl_xor(ti, off, 64);
              }
              const int i = s * H + rsub;
              const double nr = -c * ti, ni = c * tr;
              if (owner) {
                y[2 * i] = nr;
                y[2 * i + 1] = ni;
                if (i < n) {
                  vr[2 * i] += nr;
                  vr[2 * i + 1] += ni;
                  any |= (nr * nr + ni * ni > eps2) ? 1 : 0;
                }
              }
            }
            // the series' convergence test over the workgroup with one barrier: each wave
            // posts its ballot in this term's parity slot, which is next written two terms
            // later, after every thread has passed the barrier of the term in between
            const int par = k & 1;
            const bool wave_any = __any(any);
            if (lane == 0) sany[par][wave] = wave_any ? 1 : 0;
            __syncthreads();
            any = sany[par][0] | sany[par][1] | sany[par][2] | sany[par][3];
            terms++;
            double* tswap = x;
            x = y;
            y = tswap;
            if (!any || k >= kStepMaxTerms) break;
          }
        }
        for (int s = wave; s < nslices; s += 4) {
          const int i = s * H + rsub;
          if (owner && i < n) {
            vr[2 * i] *= scaling;
            vr[2 * i + 1] *= scaling;
          }
        }
        __syncthreads();
      }
      if (!p.final_state) detect(t);
    }
    if (tid == 0 && p.terms_out) p.terms_out[o] = terms;
    if (p.final_state) {
      double* ro = p.rho_out + 2 * (size_t)o * p.rho_or_stride;
      const int32_t* perm = p.sell_perm + st.perm_off;
      for (int i = tid; i < n; i += kStepThreads) {
        ro[2 * perm[i]] = vr[2 * i];
        ro[2 * perm[i] + 1] = vr[2 * i + 1];
      }
    }
    __syncthreads();
  }
}

}  // namespace spk

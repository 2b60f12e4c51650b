// Spectral evaluation of the detected trajectory for real generators.
//
// When A = dt*(H+H')/2 is real symmetric (the host checks the Q symmetry that
// guarantees it at every orientation), f(t) = coil' * exp(-i*A*t) * rho0 is
// evaluated from one eigendecomposition per orientation:
//   A = V diag(lambda) V^T,  f(t) = sum_k g_k exp(-i*lambda_k*t),
//   g_k = (coil' v_k) (v_k^T rho0).
// This is the same quantity that kernel/evolution.m:562-566 accumulates by
// stepping rho <- P*rho with P = expm(-i*L*dt) (kernel/propagator.m:40-42);
// rounding in both is bounded by t*eps*|A|, so the two agree to ~1e-12 over
// thousands of steps. Three kernels:
//   k_tridiag  Householder reduction A = Q T Q^T, T tridiagonal, with Q^T
//              applied on the fly to conj(coil) and rho0. The lower triangle
//              of the 192 x 192 FP64 matrix lives in the registers of a
//              256-thread workgroup (two per CU), so the symmetric
//              matrix-vector products and rank-2 updates of every step run
//              from registers with the Householder vectors broadcast from LDS.
//   k_dceig    eigenvalues of T and the projections g_k by divide and
//              conquer (QR on leaves of <= 6 rows, secular-equation merges
//              with one thread per root); the eigenvectors are never formed.
//   k_synth    f(64*i + j) = sum_k [g_k e^{-i lambda_k 64 i}] [e^{-i lambda_k j}]
//              as one complex matrix product per 64 x 64 block of time
//              points, on the matrix cores.
#pragma once

#include "powder_kernels.hpp"

namespace spk {

constexpr int kTriNP = 192;      // largest padded dimension of the 256-thread register layout
constexpr int kTriWNP = 256;     // largest padded dimension of the 1024-thread layout (k_tridiag_w)
constexpr int kTriGNP = 448;     // largest padded dimension with k_tridiag_w's global-memory prologue
constexpr int kSpecStride = 6 * kTriNP;  // per item: d | e | four vectors (spec_nm = kTriNP)
constexpr int kSynthB = 64;      // time block of the synthesis product

struct SpecParams {
  const SysDesc* sys;
  const Item* items;  // spectral items
  int64_t n_items;
  unsigned int* queue;  // [0] tridiagonalisation, [1] synthesis; zeroed by k_reset
  const int32_t* prow;
  const int32_t* pcol;
  const int32_t* ptpos;
  const double* vals;
  const int32_t* slots;
  const double* rho0;
  const double* rho_or;  // optional per-orientation initial states (after a dead time), stride rho_or_stride
  int64_t rho_or_stride;
  const double* coil;
  const double* alphas;
  const double* betas;
  const double* gammas;
  double dt;
  int64_t nsteps;
  double* spec;          // 6 * spec_nm doubles per spectral item: d | e | four vectors, each spec_nm long
  int spec_nm;           // kTriNP, kTriWNP or kTriGNP: the largest padded dimension class of the plan
  double* tri_work;      // k_tridiag_w's prologue: kTriGWork doubles per workgroup (plans above kTriWNP)
  int64_t tri_work_stride;
  int32_t* status;       // per item: most secular evaluations of a root, or -1 if a solve failed
  int32_t dc_max_iter;   // secular evaluations per root of the second pass (0: kDcMaxIter)
  double* per_or;        // per-orientation FIDs, interleaved complex, (nsteps+1) per orientation
};

// Entry e (row prow[e], column pcol[e] of the reduced union pattern) of the real symmetric
// generator A = dt*(x + x^T)/2 of one orientation, x = I + sum_q c_q Q_q with c_q the
// orientation's Wigner coefficients (cre/cim, indexed by the slot of each Q).
__device__ inline double real_generator_entry(const SpecParams& p, const SysDesc& sd, const double* cre,
                                              const double* cim, int e) {
  const double* vals = p.vals + 2 * sd.val_off;
  const int32_t* slots = p.slots + sd.slot_off;
  const int nnz = sd.nnz;
  const int et = p.ptpos[sd.pat_off + e];
  double xr = vals[2 * e], tr = vals[2 * et];
  for (int q = 0; q < sd.nq; q++) {
    const double* qv = vals + 2 * (size_t)(q + 1) * nnz;
    const double cr = cre[slots[q]], ci = cim[slots[q]];
    xr += cr * qv[2 * e] - ci * qv[2 * e + 1];
    tr += cr * qv[2 * et] - ci * qv[2 * et + 1];
  }
  return p.dt * 0.5 * (xr + tr);
}

// Cross-lane sums with DPP moves (VALU latency, no LDS round trip).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// sum over each aligned group of 8 lanes, returned in all 8; every lane of
// the wave must be active
__device__ __forceinline__ double sum8(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_f64<0x141>(v);  // row_half_mirror: the other quad of the 8-group
  return v;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// sum over the wave (uniform result); every lane must be active
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = sum8(v);
  v += dpp_f64<0x140>(v);  // row_mirror: the other 8-group of the 16-lane row
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// One wave forms the Householder reflection that clears row jj right of its
// first subdiagonal entry, from the row's entries x[r] at columns lane + 64 r
// and its entries xd = x_jj, x1 = x_jj+1 (uniform):
//   sigma = sum_{c >= jj+2} x_c^2,  alpha = -sign(x1) sqrt(x1^2 + sigma),
//   v = (0, ..., 0, x1 - alpha, x_{jj+2}, ...),  tau = 2/(v'v)
// (tau = 0, alpha = x1 when sigma = 0: no reflection), and writes v to vb,
// tau to *ptau and T's d_jj = xd, e_jj = alpha. Every lane of the wave active.
template <int NQ>
__device__ __forceinline__ void householder_form(const double (&x)[NQ], double xd, double x1, int jj, int lane,
                                                 double* vb, double* ptau, double* sd, double* se) {
  double part = 0.0;
#pragma unroll
  for (int r = 0; r < NQ; r++)
    if (lane + 64 * r >= jj + 2) part = fma(x[r], x[r], part);
  const double sigma = wave_sum_dpp(part);
  double tau = 0.0, alpha = x1, v1 = 0.0;
  if (sigma > 0.0) {
    const double nrm = sqrt(x1 * x1 + sigma);
    alpha = (x1 > 0.0) ? -nrm : nrm;
    v1 = x1 - alpha;
    tau = 2.0 / (v1 * v1 + sigma);
  }
#pragma unroll
  for (int r = 0; r < NQ; r++) {
    const int c = lane + 64 * r;
    vb[c] = (c > jj + 1) ? x[r] : ((c == jj + 1) ? v1 : 0.0);
  }
  if (lane == 0) {
    *ptau = tau;
    sd[jj] = xd;
    se[jj] = alpha;
  }
}

// ------------------------------------------------------------ k_tridiag
// Householder tridiagonalisation with the lower triangle of A in registers: 148 KB
// per orientation, so that two orientations (workgroups) share a CU and hide each
// other's barrier and reduction latency (the whole matrix would fill the CU's
// register file). 256 threads on a 16 x 16 grid, thread (r, c) = (tid / 16, tid % 16);
// the thread holds A(r + 16a, c + 16b) for b <= a (78 slots). The diagonal blocks
// (a == b) are held whole, above the diagonal too: the symmetric product takes their
// row sums only (each entry once), the blocks below them their row and column sums
// (each entry for itself and for its transpose), and the update and the column copy
// need no weights.
constexpr int kTriThreads = 256;
constexpr int kTriBlk = kTriNP / 16;                      // blocks of 16 rows / columns
constexpr int kTriSlots = kTriBlk * (kTriBlk + 1) / 2;  // 78
__host__ __device__ constexpr int sym_slot(int a, int b) { return a * (a + 1) / 2 + b; }

// The owners of column jj (the threads with c == jj % 16) copy its entries at rows
// >= jj (rows below jj in its block get 0 or a stale value, never read) to srow.
__device__ __forceinline__ void sym_copy_col(const double (&R)[kTriSlots], int jj, int r, int c, double* srow) {
  if (c != (jj & 15)) return;
  const int bj = jj >> 4;
#pragma unroll
  for (int b = 0; b < kTriBlk; b++) {
    if (b == bj) {
      // (the empty volatile statement keeps the branches apart: merged, they would read the
      // matrix at a run-time index, which puts the whole array in scratch memory)
      asm volatile("" ::: "memory");
#pragma unroll
      for (int a = b; a < kTriBlk; a++) srow[r + 16 * a] = R[sym_slot(a, b)];
    }
  }
}

// Column jj as it stands in the registers, reflected by wave 0 (every thread calls it).
__device__ __forceinline__ void sym_publish(const double (&R)[kTriSlots], int jj, int r, int c, int wave, int lane,
                                            double* srow, double* vb, double* ptau, double* sd, double* se) {
  sym_copy_col(R, jj, r, c, srow);
  __syncthreads();
  if (wave == 0) {
    double x[kTriNP / 64];
#pragma unroll
    for (int q = 0; q < kTriNP / 64; q++) x[q] = srow[lane + 64 * q];
    householder_form(x, srow[jj], srow[jj + 1], jj, lane, vb, ptau, sd, se);  // srow[kTriNP] stays zero
  }
}

// v (the same in every lane) as a wave-uniform value, so that branches on it are scalar
__device__ __forceinline__ double uniform_f64(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

// A*v over this thread's slots of the live column blocks (bmin <= b < bmax, the set bits of
// live): row sums into yr, column sums stored to spc[r][.]; returns this thread's part of
// v'A v from the column sums (the row sums' part is added by the caller). The row blocks of
// a live column block are not tested against bmax: rows at or beyond np hold zeros (a test
// there would be compiled to a select per product).
__device__ __forceinline__ double sym_matvec(const double (&R)[kTriSlots], const double (&vr)[kTriBlk],
                                             double (&yr)[kTriBlk], const double* v, double (*spc)[kTriNP], int r,
                                             int c, int bmin, unsigned live) {
  // the live blocks are bmin..bmax-1, so each block loads the next block's v ahead of its
  // own products
  double s = 0.0;
  double vnext = v[c + 16 * bmin];
#pragma unroll
  for (int b = 0; b < kTriBlk; b++) {
    if ((live >> b) & 1u) {
      const double vc = vnext;
      if (b + 1 < kTriBlk) vnext = v[c + 16 * (b + 1)];
      double yc = 0.0;
#pragma unroll
      for (int a = b; a < kTriBlk; a++) {  // rows at or beyond np hold zeros
        const double x = R[sym_slot(a, b)];
        yr[a] = fma(x, vc, yr[a]);
        if (a > b) yc = fma(x, vr[a], yc);  // the diagonal block gives row sums only
      }
      spc[r][c + 16 * b] = yc;
      s = fma(vc, yc, s);
    }
  }
  return s;
}

// A <- A - v w' - w v' over this thread's slots of the live column blocks (h weights the
// diagonal blocks)
__device__ __forceinline__ void sym_update(double (&R)[kTriSlots], const double (&vr)[kTriBlk],
                                           const double (&wr)[kTriBlk], const double* v, const double* sw, int c,
                                           int bmin, unsigned live) {
  double vnext = v[c + 16 * bmin], wnext = sw[c + 16 * bmin];  // (next block's loads ahead, as above)
#pragma unroll
  for (int b = 0; b < kTriBlk; b++) {
    if ((live >> b) & 1u) {
      const double vc = vnext, wc = wnext;
      if (b + 1 < kTriBlk) {
        vnext = v[c + 16 * (b + 1)];
        wnext = sw[c + 16 * (b + 1)];
      }
#pragma unroll
      for (int a = b; a < kTriBlk; a++)  // rows at or beyond np hold zeros and stay zero
        R[sym_slot(a, b)] = fma(-vr[a], wc, fma(-wr[a], vc, R[sym_slot(a, b)]));
    }
  }
}

__global__ void __launch_bounds__(kTriThreads, 2) k_tridiag(SpecParams p) {
  __shared__ double sv[2][kTriNP];  // Householder vectors, double-buffered
  __shared__ double stau[2];
  __shared__ double srow[kTriNP + 1];  // the column being reflected; srow[kTriNP] = 0
  __shared__ double sw[kTriNP];        // w = p - (tau/2)(v'p) v, p = tau*A*v
  __shared__ double spr[2][kTriNP];    // row sums of A*v over column groups 0-7 and 8-15
  __shared__ double spc[16][kTriNP];   // column sums of A*v, one row per row group r
  __shared__ double sq[kTriThreads / 64];
  __shared__ double su[4][kTriNP];
  __shared__ double sd[kTriNP], se[kTriNP];
  __shared__ double cre[34], cim[34];
  __shared__ int sitem;
  // 16 rows of the generator under assembly (rows 16a..16a+15 in pass a); all zero between uses
  constexpr int kAsmStride = kTriNP + 4;
  __shared__ double sA[16 * kAsmStride];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
  const int r = tid >> 4, c = tid & 15;
  for (int k = tid; k < 16 * kAsmStride; k += kTriThreads) sA[k] = 0.0;
  for (int k = tid; k <= kTriNP; k += kTriThreads) srow[k] = 0.0;

  for (;;) {
    if (tid == 0) sitem = (int)atomicAdd(&p.queue[0], 1u);
    __syncthreads();
    const int64_t it = __builtin_amdgcn_readfirstlane(sitem);  // uniform (scalar branches below)
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sdesc = p.sys[item.sys];
    const int n = __builtin_amdgcn_readfirstlane(sdesc.n), np = __builtin_amdgcn_readfirstlane(sdesc.np);
    const int bmax = (np + 15) >> 4;  // blocks below the padded dimension
    const int64_t o = item.orient;
    if (tid < 34) wigner_coeff_slot(p.alphas[o], p.betas[o], p.gammas[o], tid, cre, cim);
    {
      const double* coilv = p.coil + 2 * sdesc.vec_off;
      const double* rhov = p.rho_or ? p.rho_or + 2 * (size_t)o * p.rho_or_stride : p.rho0 + 2 * sdesc.vec_off;
      for (int i = tid; i < kTriNP; i += kTriThreads) {
        double c0 = 0.0, c1 = 0.0, r0 = 0.0, r1 = 0.0;
        if (i < n) {
          c0 = coilv[2 * i];
          c1 = -coilv[2 * i + 1];
          r0 = rhov[2 * i];
          r1 = rhov[2 * i + 1];
        }
        su[0][i] = c0;
        su[1][i] = c1;
        su[2][i] = r0;
        su[3][i] = r1;
        sd[i] = 0.0;
        se[i] = 0.0;
      }
    }
    __syncthreads();
    // the lower triangle, 16 rows per pass: the pass's lower pattern entries are
    // scattered into sA, then every thread takes its slots of row r + 16a and clears them.
    // The values of the first kAsmCache * 256 pattern entries are computed up front (their
    // loads all in flight at once), entry tid + 256k in this thread's registers with its
    // position (row << 8 | column, -1 for an upper entry); later entries are read per pass.
    constexpr int kAsmCache = 8;
    double R[kTriSlots];
    const int32_t* prow = p.prow + sdesc.pat_off;
    const int32_t* pcol = p.pcol + sdesc.pat_off;
    const int nnz = sdesc.nnz;
    double cval[kAsmCache];
    int cpos[kAsmCache];
#pragma unroll
    for (int k = 0; k < kAsmCache; k++) {
      const int e = tid + kTriThreads * k;
      const int ee = (e < nnz) ? e : 0;
      const int i = prow[ee], jc = pcol[ee];
      cval[k] = (e < nnz) ? real_generator_entry(p, sdesc, cre, cim, ee) : 0.0;
      cpos[k] = (e < nnz && (jc >> 4) <= (i >> 4)) ? ((i << 8) | jc) : -1;
    }
#pragma unroll
    for (int a = 0; a < kTriBlk; a++) {
      if (a < bmax) {
#pragma unroll
        for (int k = 0; k < kAsmCache; k++)
          if (cpos[k] >= 0 && (cpos[k] >> 12) == a)
            sA[((cpos[k] >> 8) - 16 * a) * kAsmStride + (cpos[k] & 255)] = cval[k];
        for (int e = max(sdesc.rb[a], kAsmCache * kTriThreads) + tid; e < sdesc.rb[a + 1]; e += kTriThreads) {
          const int i = prow[e], jc = pcol[e];
          if ((jc >> 4) <= (i >> 4)) sA[(i - 16 * a) * kAsmStride + jc] = real_generator_entry(p, sdesc, cre, cim, e);
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b <= a; b++) {
          double* s = &sA[r * kAsmStride + c + 16 * b];
          R[sym_slot(a, b)] = *s;
          *s = 0.0;
        }
        __syncthreads();
      } else {
#pragma unroll
        for (int b = 0; b <= a; b++) R[sym_slot(a, b)] = 0.0;
      }
    }
    sym_publish(R, 0, r, c, wave, lane, srow, sv[0], &stau[0], sd, se);
    __syncthreads();

    for (int j = 0; j < n - 2; j++) {
      const int cb = j & 1;
      const double* v = sv[cb];
      const double tau = uniform_f64(stau[cb]);  // reflection of column j
      const int jn = j + 1;         // the next column to reflect
      const int bmin = jn >> 4;     // blocks below bmin are eliminated
      // the live column blocks bmin..bmax-1 as a bit mask (one bit test per block)
      const unsigned live = ((1u << bmax) - 1u) & ~((1u << bmin) - 1u);
      if (tau != 0.0) {             // uniform
        // the wave that reflects column jn rotates with the step, so that the two workgroups
        // sharing a CU rarely have their look-ahead waves on the same SIMD
        const int la = jn & 3;
        sym_copy_col(R, jn, r, c, srow);  // column jn as the previous step left it
        // A*v: row sums yr over this thread's columns, column sums yc over its rows
        double vr[kTriBlk], yr[kTriBlk];
#pragma unroll
        for (int a = 0; a < kTriBlk; a++) {
          vr[a] = v[r + 16 * a];
          yr[a] = 0.0;
        }
        // this thread's part of v'A v
        double s = sym_matvec(R, vr, yr, v, spc, r, c, bmin, live);
        // the reductions that end the phase are latency-bound: raised priority (4.61 -> 4.56 ms)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int a = 0; a < kTriBlk; a++) {
          if ((live >> a) & 1u) {
            s = fma(vr[a], yr[a], s);
            const double t = sum8(yr[a]);  // (the last 8-group sum is left to the assembly)
            if ((c & 7) == 0) spr[c >> 3][r + 16 * a] = t;
          }
        }
        const double qw = wave_sum_dpp(s) * tau;
        if (lane == 0) sq[wave] = qw;
        __syncthreads();
        // the assembly, a short latency-bound phase, goes ahead of the other workgroup's waves
        // (4.64 -> 4.61 ms)
        __builtin_amdgcn_s_setprio(2);
        // A <- A - v w' - w v' with w = p - (tau/2)(v'p) v, assembled with p
        if (tid < kTriNP) {
          const double K = 0.5 * tau * ((sq[0] + sq[1]) + (sq[2] + sq[3]));
          double t4[4];
#pragma unroll
          for (int q = 0; q < 4; q++) t4[q] = (spc[q][tid] + spc[q + 4][tid]) + (spc[q + 8][tid] + spc[q + 12][tid]);
          const double t = (spr[0][tid] + spr[1][tid]) + ((t4[0] + t4[1]) + (t4[2] + t4[3]));
          const double pk = (tid > j && tid < np) ? t * tau : 0.0;
          sw[tid] = pk - K * v[tid];
        }
        __syncthreads();
        // from here to the end of the step (the look-ahead, Q^T and the update) the workgroup's
        // waves run at raised priority over the product phase of the CU's other workgroup, the
        // reflecting wave highest (measured: 4.83 -> 4.65 ms against raising the reflecting
        // wave alone for the whole step)
        __builtin_amdgcn_s_setprio(1);
        if (wave == la) __builtin_amdgcn_s_setprio(3);
        if (wave == la) {
          // column jn after this update, from its copy, and its reflection for the next step
          const double vn = v[jn], wn = sw[jn];
          double x[kTriNP / 64];
#pragma unroll
          for (int q = 0; q < kTriNP / 64; q++) {
            const int i = lane + 64 * q;
            const double vi = v[i], wi = sw[i];
            x[q] = fma(-vn, wi, fma(-wn, vi, srow[i]));
          }
          const double xd = fma(-vn, wn, fma(-wn, vn, srow[jn]));
          const double v1 = v[jn + 1], w1 = sw[jn + 1];
          const double x1 = fma(-vn, w1, fma(-wn, v1, srow[jn + 1]));
          householder_form(x, xd, x1, jn, lane, sv[cb ^ 1], &stau[cb ^ 1], sd, se);
        } else {
          // Q^T applied to conj(coil) and rho0: the first wave after la takes vectors 0 and 3,
          // the other two one each
          const int qw0 = (wave - la - 1) & 3;
#pragma unroll
          for (int pass = 0; pass < 2; pass++) {
            const int q = (pass == 0) ? qw0 : 3;
            if (pass == 0 || qw0 == 0) {
              double vv[kTriNP / 64], uu[kTriNP / 64];
              double t = 0.0;
#pragma unroll
              for (int k = 0; k < kTriNP / 64; k++) {
                const int i = lane + 64 * k;
                vv[k] = v[i];
                uu[k] = su[q][i];
                t = fma(vv[k], uu[k], t);
              }
              t = wave_sum_dpp(t) * tau;
#pragma unroll
              for (int k = 0; k < kTriNP / 64; k++)
                if (vv[k] != 0.0) su[q][lane + 64 * k] = uu[k] - t * vv[k];
            }
          }
        }
        double wr[kTriBlk];
#pragma unroll
        for (int a = 0; a < kTriBlk; a++) wr[a] = sw[r + 16 * a];
        sym_update(R, vr, wr, v, sw, c, bmin, live);
        __builtin_amdgcn_s_setprio(0);
      } else {
        // no reflection at this step: column jn stands as it is in the registers
        sym_publish(R, jn, r, c, wave, lane, srow, sv[cb ^ 1], &stau[cb ^ 1], sd, se);
      }
      __syncthreads();
    }
    if (n >= 2) sym_publish(R, n - 1, r, c, wave, lane, srow, sv[(n - 1) & 1], &stau[(n - 1) & 1], sd, se);
    __syncthreads();
    const int nm = p.spec_nm;
    double* out = p.spec + (size_t)it * 6 * nm;
    for (int i = tid; i < kTriNP; i += kTriThreads) {
      out[i] = sd[i];
      out[nm + i] = se[i];
#pragma unroll
      for (int q = 0; q < 4; q++) out[(2 + q) * nm + i] = su[q][i];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ k_tridiag_w
// The same Householder tridiagonalisation for padded dimensions kTriNP < np <= kTriWNP: the
// lower triangle of the 256 x 256 FP64 matrix lives in the registers of one 512-thread
// workgroup per CU (two waves per SIMD, so that each thread may hold 144 VGPRs of it without
// spilling). Thread (r, c) = (tid / 32, tid % 32) holds A(r + 16a, c + 32b) for the blocks of
// 16 rows and 32 columns that reach the lower triangle, a >= 2b (72 slots). The two blocks of
// column block b that make up its 32 x 32 diagonal square (a = 2b, 2b + 1) hold the whole
// square, above the diagonal too: the product takes their row sums only (each entry of the
// square once), the blocks below the square their row and column sums (each entry for itself
// and for its transpose), and the update and the column copies need no weights. Wave w holds
// rows r = 2w (lanes 0-31) and 2w + 1 (lanes 32-63). Row sums of A*v are reduced over each
// 16-lane DPP row (two partial sums per index), column sums over the wave's two rows by a lane
// exchange (one per wave and index).
constexpr int kTriWThreads = 512;
constexpr int kTriWWaves = kTriWThreads / 64;  // 8
constexpr int kTriWRB = kTriWNP / 16;          // 16 row blocks of 16
constexpr int kTriWCB = kTriWNP / 32;          // 8 column blocks of 32
constexpr int kTriWSlots = kTriWCB * (kTriWCB + 1);  // 72: column block b has rows blocks 2b .. 15
// slot of row block a (>= 2b) of column block b
__host__ __device__ constexpr int symw_slot(int a, int b) { return b * (2 * kTriWCB + 1 - b) + (a - 2 * b); }

// sum over each 16-lane DPP row, returned in all 16 lanes; every lane of the wave active
__device__ __forceinline__ double sum16(double v) {
  v = sum8(v);
  v += dpp_f64<0x140>(v);  // row_mirror: the other 8-group of the row
  return v;
}

// The owners of column jj (c == jj % 32) copy its entries at rows >= 32 (jj / 32) to srow (rows
// below jj are never read).
__device__ __forceinline__ void symw_copy_col(const double (&R)[kTriWSlots], int jj, int r, int c, double* srow) {
  if (c != (jj & 31)) return;
  const int bj = jj >> 5;
#pragma unroll
  for (int b = 0; b < kTriWCB; b++) {
    if (b == bj) {
      // (the empty volatile statement keeps the eight branches apart: merged, they would read
      // the matrix at a run-time index, which puts the whole array in scratch memory)
      asm volatile("" ::: "memory");
#pragma unroll
      for (int a = 2 * b; a < kTriWRB; a++) srow[r + 16 * a] = R[symw_slot(a, b)];
    }
  }
}

// Column jj as it stands in the registers, reflected by wave 0 (every thread calls it).
__device__ __forceinline__ void symw_publish(const double (&R)[kTriWSlots], int jj, int r, int c, int wave, int lane,
                                             double* srow, double* vb, double* ptau, double* sd, double* se) {
  symw_copy_col(R, jj, r, c, srow);
  __syncthreads();
  if (wave == 0) {
    double x[kTriWNP / 64];
#pragma unroll
    for (int q = 0; q < kTriWNP / 64; q++) x[q] = srow[lane + 64 * q];
    householder_form(x, srow[jj], srow[jj + 1], jj, lane, vb, ptau, sd, se);
  }
}

// A*v over this thread's slots of the live column blocks (b >= bmin), in two passes over the
// registers (one row-block-indexed array live at a time beside the matrix): row sums, reduced
// over each 16-lane row (two partial sums per index, to spr), then column sums of the blocks
// below the diagonal squares, the wave's two rows added (one partial sum per wave and index, to
// spc). Returns this thread's part of v'A v. Rows and columns at or beyond np hold zeros.
__device__ __forceinline__ double symw_matvec(const double (&R)[kTriWSlots], const double* v, double (*spr)[kTriWNP],
                                              double (*spc)[kTriWNP], int wave, int lane, int r, int c, int bmin,
                                              unsigned live) {
  double s = 0.0;
  {
    double yr[kTriWRB];
#pragma unroll
    for (int a = 0; a < kTriWRB; a++) yr[a] = 0.0;
    double vnext = v[c + 32 * bmin];
#pragma unroll
    for (int b = 0; b < kTriWCB; b++) {
      if ((live >> b) & 1u) {
        const double vc = vnext;
        if (b + 1 < kTriWCB) vnext = v[c + 32 * (b + 1)];
#pragma unroll
        for (int a = 2 * b; a < kTriWRB; a++) yr[a] = fma(R[symw_slot(a, b)], vc, yr[a]);
      }
    }
#pragma unroll
    for (int a = 0; a < kTriWRB; a++) {
      if ((live >> (a >> 1)) & 1u) {  // row block a meets the live column blocks bmin .. a / 2
        s = fma(v[r + 16 * a], yr[a], s);
        const double t = sum16(yr[a]);
        if ((lane & 15) == 0) spr[(lane >> 4) & 1][r + 16 * a] = t;
      }
    }
  }
  double vr[kTriWRB];
#pragma unroll
  for (int a = 0; a < kTriWRB; a++) vr[a] = v[r + 16 * a];
#pragma unroll
  for (int b = 0; b < kTriWCB; b++) {
    if ((live >> b) & 1u) {
      double yc = 0.0;  // below the diagonal square only
#pragma unroll
      for (int a = 2 * b + 2; a < kTriWRB; a++) yc = fma(R[symw_slot(a, b)], vr[a], yc);
      s = fma(v[c + 32 * b], yc, s);
      const double yo = __shfl_xor(yc, 32);  // the wave's other row, same column
      if (lane < 32) spc[wave][c + 32 * b] = yc + yo;
    }
  }
  return s;
}

// A <- A - v w' - w v' over this thread's slots of the live column blocks, for the row blocks
// A0 .. A0 + 7 (two calls cover the sixteen: half the row values are live at a time)
template <int A0>
__device__ __forceinline__ void symw_update(double (&R)[kTriWSlots], const double* v, const double* sw, int r, int c,
                                            int bmin, unsigned live) {
  double vr[8], wr[8];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    vr[a] = v[r + 16 * (A0 + a)];
    wr[a] = sw[r + 16 * (A0 + a)];
  }
  double vnext = v[c + 32 * bmin], wnext = sw[c + 32 * bmin];
#pragma unroll
  for (int b = 0; 2 * b < A0 + 8; b++) {
    if ((live >> b) & 1u) {
      const double vc = vnext, wc = wnext;
      if (2 * (b + 1) < A0 + 8) {
        vnext = v[c + 32 * (b + 1)];
        wnext = sw[c + 32 * (b + 1)];
      }
#pragma unroll
      for (int a = (2 * b > A0 ? 2 * b : A0); a < A0 + 8; a++)
        R[symw_slot(a, b)] = fma(-vr[a - A0], wc, fma(-wr[a - A0], vc, R[symw_slot(a, b)]));
    }
  }
}

// Global-memory prologue of k_tridiag_w for padded dimensions kTriWNP < np <= kTriGNP: the
// first k0 = np - kTriWNP (rounded up to a multiple of 32) columns are reduced with the matrix in
// the workgroup's slot of a global workspace, after which the trailing block (at most kTriWNP
// rows) is loaded into the register layout and the reduction continues there. The lower
// triangle is stored as 32 x 32 blocks (block (a, b), b <= a, at sym_slot(a, b) * 1024, row-major
// within the block), h-weighted in the diagonal blocks as the register layout is, so the
// hand-off is a copy. Each step makes one pass over the live blocks, which applies the previous
// step's rank-2 update, stores the block and forms the next symmetric product from the updated
// values (the next reflection comes first, from its column updated in LDS). The pass streams
// the trailing lower triangle once from and to HBM per step: it is bound by that traffic.
constexpr int kTriGBlk = kTriGNP / 32;                          // 14 block rows
constexpr int kTriGBlocks = kTriGBlk * (kTriGBlk + 1) / 2;      // 105 stored blocks
constexpr int64_t kTriGWork = (int64_t)kTriGBlocks * 1024;      // workspace doubles per workgroup

// The pass over the live blocks (block rows and columns >= bmin, below nbk), split into equal
// runs of consecutive blocks (row-major) over the waves. Lane (c, h2) = (lane % 32, lane / 32)
// takes column c and rows 16 h2 .. 16 h2 + 15 of each block; the next block's values are loaded
// while the current one is processed. When upd: A <- A - v w' - w v' (vo, wo) and the block is
// stored back; when prod: row sums of A*vn (reduced over the 32 lanes of each half-wave, one
// per row index and wave) to rowp, column sums (the two halves added) to colp, and the lane's
// part of vn'A vn is returned. rowp and colp are zero beforehand.
__device__ __forceinline__ double trig_pass(double* W, int nbk, int bmin, bool upd, bool prod, const double* vo,
                                            const double* wo, const double* vn, double (*rowp)[kTriGNP],
                                            double (*colp)[kTriGNP], int wave, int lane) {
  const int m = nbk - bmin;
  const int L = m * (m + 1) / 2;
  const int f0 = (L * wave) / kTriWWaves, f1 = (L * (wave + 1)) / kTriWWaves;
  const int c = lane & 31, h2 = lane >> 5;
  double s = 0.0;
  if (f0 >= f1) return s;
  int ap = 0;
  while ((ap + 1) * (ap + 2) / 2 <= f0) ap++;
  int bp = f0 - ap * (ap + 1) / 2;
  double racc[16], xn[16];
#pragma unroll
  for (int i = 0; i < 16; i++) racc[i] = 0.0;
  {
    const double* nb = W + (size_t)sym_slot(bmin + ap, bmin + bp) * 1024 + 512 * h2 + c;
#pragma unroll
    for (int i = 0; i < 16; i++) xn[i] = nb[32 * i];
  }
  for (int f = f0; f < f1; f++) {
    const int a = bmin + ap, b = bmin + bp;
    double* blk = W + (size_t)sym_slot(a, b) * 1024 + 512 * h2 + c;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = xn[i];
    const bool row_end = (bp == ap) || (f + 1 == f1);
    if (++bp > ap) {
      ap++;
      bp = 0;
    }
    if (f + 1 < f1) {
      const double* nb = W + (size_t)sym_slot(bmin + ap, bmin + bp) * 1024 + 512 * h2 + c;
#pragma unroll
      for (int i = 0; i < 16; i++) xn[i] = nb[32 * i];
    }
    const int C = 32 * b + c, R0 = 32 * a + 16 * h2;
    const double vC = vo[C], wC = wo[C], nC = vn[C];
    if (upd) {
      if (a == b) {  // h-weighted diagonal block
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const int r = 16 * h2 + i;
          const double hh = (r > c) ? 1.0 : ((r == c) ? 0.5 : 0.0);
          x[i] = fma(-hh * vo[R0 + i], wC, fma(-hh * wo[R0 + i], vC, x[i]));
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; i++) x[i] = fma(-vo[R0 + i], wC, fma(-wo[R0 + i], vC, x[i]));
      }
#pragma unroll
      for (int i = 0; i < 16; i++) blk[32 * i] = x[i];
    }
    if (prod) {
      double cacc = 0.0;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        racc[i] = fma(x[i], nC, racc[i]);
        cacc = fma(x[i], vn[R0 + i], cacc);
      }
      s = fma(nC, cacc, s);
      cacc += __shfl_xor(cacc, 32);  // rows of the other half, same column
      if (h2 == 0) colp[wave][C] += cacc;
      // the block row's sums are complete when the run leaves it
      if (row_end) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
          double t = sum16(racc[i]);
          t += __shfl_xor(t, 16);  // the other 16 columns of the half
          racc[i] = t;
        }
        if (c == 0) {
#pragma unroll
          for (int i = 0; i < 16; i++) rowp[wave][R0 + i] = racc[i];
        }
#pragma unroll
        for (int i = 0; i < 16; i++) racc[i] = 0.0;
      }
    }
  }
  return 2.0 * s;  // the row sums' half of vn'A vn equals the column sums' half
}

__global__ void __launch_bounds__(kTriWThreads, 1) k_tridiag_w(SpecParams p) {
  __shared__ double sv[2][kTriGNP];  // Householder vectors, double-buffered
  __shared__ double stau[2];
  __shared__ double srow[kTriGNP + 1];  // the column being reflected; zero beyond np
  __shared__ double sw[kTriGNP];        // w = p - (tau/2)(v'p) v, p = tau*A*v
  __shared__ double spr[2][kTriWNP];    // row sums of A*v over columns c < 16 and c >= 16
  __shared__ double sq[kTriWWaves];
  __shared__ double su[4][kTriGNP];
  __shared__ double sd[kTriGNP], se[kTriGNP];
  __shared__ double cre[34], cim[34];
  __shared__ int sitem;
  // one region, used in turn by the assembly tile and the column sums of the register layout and
  // by the prologue's partial sums
  constexpr int kAsmStride = kTriWNP + 4;
  __shared__ double lds_u[2 * kTriWWaves * kTriGNP];
  static_assert(16 * kAsmStride <= 2 * kTriWWaves * kTriGNP, "assembly tile");
  double* sA = lds_u;  // 16 rows of the generator under assembly
  double (*spc)[kTriWNP] = reinterpret_cast<double (*)[kTriWNP]>(lds_u);  // column sums, one row per wave
  double (*rowp)[kTriGNP] = reinterpret_cast<double (*)[kTriGNP]>(lds_u);
  double (*colp)[kTriGNP] = reinterpret_cast<double (*)[kTriGNP]>(lds_u + kTriWWaves * kTriGNP);
  const int tid = threadIdx.x;
  // the wave index as a wave-uniform value: the prologue's block runs and the reflecting wave's
  // branches are then scalar
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = tid >> 5, c = tid & 31;  // r < 16
  for (int k = tid; k <= kTriGNP; k += kTriWThreads) srow[k] = 0.0;
  double* W = p.tri_work ? p.tri_work + (size_t)blockIdx.x * p.tri_work_stride : nullptr;

  for (;;) {
    if (tid == 0) sitem = (int)atomicAdd(&p.queue[0], 1u);
    __syncthreads();
    const int64_t it = __builtin_amdgcn_readfirstlane(sitem);  // uniform (scalar branches below)
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sdesc = p.sys[item.sys];
    const int n = __builtin_amdgcn_readfirstlane(sdesc.n), np = __builtin_amdgcn_readfirstlane(sdesc.np);
    // columns reduced by the prologue (a multiple of 32, so the trailing block is block-aligned)
    const int k0 = (np > kTriWNP) ? ((np - kTriWNP + 31) >> 5) << 5 : 0;
    const int n2 = n - k0, np2 = np - k0;
    const int cmax = (np2 + 31) >> 5;  // column blocks below the padded dimension of the register stage
    const int64_t o = item.orient;
    if (tid < 34) wigner_coeff_slot(p.alphas[o], p.betas[o], p.gammas[o], tid, cre, cim);
    if (tid < kTriGNP) {
      const double* coilv = p.coil + 2 * sdesc.vec_off;
      const double* rhov = p.rho_or ? p.rho_or + 2 * (size_t)o * p.rho_or_stride : p.rho0 + 2 * sdesc.vec_off;
      const int i = tid;
      double c0 = 0.0, c1 = 0.0, r0 = 0.0, r1 = 0.0;
      if (i < n) {
        c0 = coilv[2 * i];
        c1 = -coilv[2 * i + 1];
        r0 = rhov[2 * i];
        r1 = rhov[2 * i + 1];
      }
      su[0][i] = c0;
      su[1][i] = c1;
      su[2][i] = r0;
      su[3][i] = r1;
      sd[i] = 0.0;
      se[i] = 0.0;
    }
    const int32_t* prow = p.prow + sdesc.pat_off;
    const int32_t* pcol = p.pcol + sdesc.pat_off;
    double R[kTriWSlots];
    if (k0 > 0) {
      // ---- prologue: the lower triangle in the global workspace, h-weighted diagonal blocks
      const int nbk = (np + 31) >> 5;
      const int nblk = nbk * (nbk + 1) / 2;
      for (int e = tid; e < nblk * 1024; e += kTriWThreads) W[e] = 0.0;
      __syncthreads();
      for (int e = tid; e < sdesc.nnz; e += kTriWThreads) {
        const int i = prow[e], jc = pcol[e];
        if (jc <= i) {
          const double x = real_generator_entry(p, sdesc, cre, cim, e);
          W[(size_t)sym_slot(i >> 5, jc >> 5) * 1024 + (i & 31) * 32 + (jc & 31)] = (i == jc) ? 0.5 * x : x;
        }
      }
      __syncthreads();
      // column 0 and its reflection, then A*v0 (no update); the column is zero beyond np (the
      // reflection sums it to kTriGNP, and an earlier item may have left values there)
      if (tid < kTriGNP)
        srow[tid] = (tid < np) ? W[(size_t)sym_slot(tid >> 5, 0) * 1024 + (tid & 31) * 32] * (tid == 0 ? 2.0 : 1.0)
                               : 0.0;
      for (int k = tid; k < 2 * kTriWWaves * kTriGNP; k += kTriWThreads) lds_u[k] = 0.0;
      __syncthreads();
      if (wave == 0) {
        double x[kTriGNP / 64];
#pragma unroll
        for (int q = 0; q < kTriGNP / 64; q++) x[q] = srow[lane + 64 * q];
        householder_form(x, srow[0], srow[1], 0, lane, sv[0], &stau[0], sd, se);
      }
      __syncthreads();
      {
        const double s = trig_pass(W, nbk, 0, false, true, sv[0], sw, sv[0], rowp, colp, wave, lane);
        const double q = wave_sum_dpp(s);
        if (lane == 0) sq[wave] = q;
      }
      __syncthreads();
      for (int j = 0; j < k0; j++) {
        const int cb = j & 1;
        const double* v = sv[cb];
        const double tau = uniform_f64(stau[cb]);
        // w for reflection j from the partial sums of A*v (rows and columns > j)
        if (tid < kTriGNP) {
          double qs = 0.0, t = 0.0;
#pragma unroll
          for (int w = 0; w < kTriWWaves; w++) {
            qs += sq[w];
            t += rowp[w][tid] + colp[w][tid];
          }
          const double pk = (tid > j && tid < np) ? t * tau : 0.0;
          sw[tid] = pk - 0.5 * tau * (qs * tau) * v[tid];
        }
        __syncthreads();
        const bool nxt = (j + 1 < k0);
        if (wave >= 1 && wave <= 4 && tau != 0.0) {
          // Q^T applied to conj(coil) and rho0 (real and imaginary parts), one vector per wave
          const int q = wave - 1;
          double vv[kTriGNP / 64], uu[kTriGNP / 64];
          double t = 0.0;
#pragma unroll
          for (int k = 0; k < kTriGNP / 64; k++) {
            const int i = lane + 64 * k;
            vv[k] = v[i];
            uu[k] = su[q][i];
            t = fma(vv[k], uu[k], t);
          }
          t = wave_sum_dpp(t) * tau;
#pragma unroll
          for (int k = 0; k < kTriGNP / 64; k++)
            if (vv[k] != 0.0) su[q][lane + 64 * k] = uu[k] - t * vv[k];
        }
        if (nxt) {
          // column j + 1 as stored (before this step's update), and the partial sums cleared
          const int jn = j + 1;
          if (tid >= jn && tid < kTriGNP)
            srow[tid] = (tid < np) ? W[(size_t)sym_slot(tid >> 5, jn >> 5) * 1024 + (tid & 31) * 32 + (jn & 31)] *
                                         (tid == jn ? 2.0 : 1.0)
                                   : 0.0;
          for (int k = tid; k < 2 * kTriWWaves * kTriGNP; k += kTriWThreads) lds_u[k] = 0.0;
          __syncthreads();
          if (wave == 0) {
            // the column after this step's update, and its reflection
            const double vn = v[jn], wn = sw[jn];
            double x[kTriGNP / 64];
#pragma unroll
            for (int q = 0; q < kTriGNP / 64; q++) {
              const int i = lane + 64 * q;
              x[q] = fma(-vn, sw[i], fma(-wn, v[i], srow[i]));
            }
            const double xd = fma(-vn, wn, fma(-wn, vn, srow[jn]));
            const double x1 = fma(-vn, sw[jn + 1], fma(-wn, v[jn + 1], srow[jn + 1]));
            householder_form(x, xd, x1, jn, lane, sv[cb ^ 1], &stau[cb ^ 1], sd, se);
          }
          __syncthreads();
        }
        if (tau != 0.0 || nxt) {
          const double s = trig_pass(W, nbk, (j + 1) >> 5, tau != 0.0, nxt, v, sw, sv[cb ^ 1], rowp, colp, wave,
                                     lane);
          if (nxt) {
            const double q = wave_sum_dpp(s);
            if (lane == 0) sq[wave] = q;
          }
        }
        __syncthreads();
      }
      // ---- hand-off: the trailing block (rows and columns >= k0) into the register layout;
      // local row block a (16 rows) lies in stored block row (k0 + 16a) / 32, half a % 2; the
      // diagonal squares are filled above the diagonal from the transposed entries and their
      // diagonal restored from the stored halves
      const int ob = k0 >> 5;
#pragma unroll
      for (int b = 0; b < kTriWCB; b++) {
#pragma unroll
        for (int a = 2 * b; a < kTriWRB; a++) {
          const int rr = 16 * (a & 1) + r;  // row within the stored 32 x 32 block
          const double* blk = W + (size_t)sym_slot((a >> 1) + ob, b + ob) * 1024;
          double x = 0.0;
          if (b < cmax) {
            if (a >= 2 * b + 2 || rr > c)
              x = blk[rr * 32 + c];
            else if (rr == c)
              x = 2.0 * blk[rr * 32 + c];
            else
              x = blk[c * 32 + rr];
          }
          R[symw_slot(a, b)] = x;
        }
      }
      __syncthreads();
    } else {
      for (int k = tid; k < 16 * kAsmStride; k += kTriWThreads) sA[k] = 0.0;
      __syncthreads();
      // the lower triangle and the diagonal squares, one row block (16 rows) per pass: the pass's
      // pattern entries (rb[a] .. rb[a + 1]) at or below the diagonal square are scattered into
      // sA, then every thread takes its slots of row r + 16a and clears them
#pragma unroll
      for (int a = 0; a < kTriWRB; a++) {
        if (a < 2 * cmax) {
          const int e1 = sdesc.rb[a + 1];
          for (int e = sdesc.rb[a] + tid; e < e1; e += kTriWThreads) {
            const int i = prow[e], jc = pcol[e];
            if ((jc >> 5) <= (i >> 5)) sA[(i - 16 * a) * kAsmStride + jc] = real_generator_entry(p, sdesc, cre, cim, e);
          }
          __syncthreads();
#pragma unroll
          for (int b = 0; 2 * b <= a; b++) {
            double* s = &sA[r * kAsmStride + c + 32 * b];
            R[symw_slot(a, b)] = *s;
            *s = 0.0;
          }
          __syncthreads();
        } else {
#pragma unroll
          for (int b = 0; 2 * b <= a; b++) R[symw_slot(a, b)] = 0.0;
        }
      }
    }
    // ---- register stage on rows and columns k0 .. np - 1 (local index i = global i - k0); the
    // offset vectors are formed by address arithmetic (an array of pointers indexed at run time
    // would live in scratch memory)
    double* srowo = srow + k0;
    double* swo = sw + k0;
    double* sdo = sd + k0;
    double* seo = se + k0;
    symw_publish(R, 0, r, c, wave, lane, srowo, &sv[0][k0], &stau[0], sdo, seo);
    __syncthreads();

    for (int j = 0; j < n2 - 2; j++) {
      const int cb = j & 1;
      const double* v = &sv[cb][k0];
      const double tau = uniform_f64(stau[cb]);  // reflection of column j
      const int jn = j + 1;                      // the next column to reflect
      const int bmin = jn >> 5;                  // column blocks below bmin are eliminated
      const unsigned live = ((1u << cmax) - 1u) & ~((1u << bmin) - 1u);
      if (tau != 0.0) {  // uniform
        const int la = jn & (kTriWWaves - 1);  // the wave that reflects column jn
        symw_copy_col(R, jn, r, c, srowo);  // column jn as the previous step left it
        const double s = symw_matvec(R, v, spr, spc, wave, lane, r, c, bmin, live);
        __builtin_amdgcn_s_setprio(1);
        const double qw = wave_sum_dpp(s) * tau;
        if (lane == 0) sq[wave] = qw;
        __syncthreads();
        __builtin_amdgcn_s_setprio(2);
        // A <- A - v w' - w v' with w = p - (tau/2)(v'p) v, assembled with p
        if (tid < kTriWNP) {
          double t4[4];
#pragma unroll
          for (int q = 0; q < 4; q++) t4[q] = spc[q][tid] + spc[q + 4][tid];
          const double K = 0.5 * tau * (((sq[0] + sq[4]) + (sq[1] + sq[5])) + ((sq[2] + sq[6]) + (sq[3] + sq[7])));
          const double t = (spr[0][tid] + spr[1][tid]) + ((t4[0] + t4[1]) + (t4[2] + t4[3]));
          const double pk = (tid > j && tid < np2) ? t * tau : 0.0;
          swo[tid] = pk - K * v[tid];
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);
        if (wave == la) {
          __builtin_amdgcn_s_setprio(3);
          // column jn after this update, from its copy, and its reflection for the next step
          const double vn = v[jn], wn = swo[jn];
          double x[kTriWNP / 64];
#pragma unroll
          for (int q = 0; q < kTriWNP / 64; q++) {
            const int i = lane + 64 * q;
            x[q] = fma(-vn, swo[i], fma(-wn, v[i], srowo[i]));
          }
          const double xd = fma(-vn, wn, fma(-wn, vn, srowo[jn]));
          const double x1 = fma(-vn, swo[jn + 1], fma(-wn, v[jn + 1], srowo[jn + 1]));
          householder_form(x, xd, x1, jn, lane, &sv[cb ^ 1][k0], &stau[cb ^ 1], sdo, seo);
        } else {
          // Q^T applied to conj(coil) and rho0 (real and imaginary parts): the four waves
          // after la take one vector each
          const int q = ((wave - la) & (kTriWWaves - 1)) - 1;
          if (q < 4) {
            double* uq = &su[q][k0];
            double vv[kTriWNP / 64], uu[kTriWNP / 64];
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < kTriWNP / 64; k++) {
              const int i = lane + 64 * k;
              vv[k] = v[i];
              uu[k] = uq[i];
              t = fma(vv[k], uu[k], t);
            }
            t = wave_sum_dpp(t) * tau;
#pragma unroll
            for (int k = 0; k < kTriWNP / 64; k++)
              if (vv[k] != 0.0) uq[lane + 64 * k] = uu[k] - t * vv[k];
          }
        }
        symw_update<0>(R, v, swo, r, c, bmin, live);
        symw_update<8>(R, v, swo, r, c, bmin, live);
        __builtin_amdgcn_s_setprio(0);
      } else {
        // no reflection at this step: column jn stands as it is in the registers
        symw_publish(R, jn, r, c, wave, lane, srowo, &sv[cb ^ 1][k0], &stau[cb ^ 1], sdo, seo);
      }
      __syncthreads();
    }
    if (n2 >= 2)
      symw_publish(R, n2 - 1, r, c, wave, lane, srowo, &sv[(n2 - 1) & 1][k0], &stau[(n2 - 1) & 1], sdo, seo);
    __syncthreads();
    const int nm = p.spec_nm;  // kTriWNP or kTriGNP
    double* out = p.spec + (size_t)it * 6 * nm;
    for (int i = tid; i < nm; i += kTriWThreads) {
      out[i] = sd[i];
      out[nm + i] = se[i];
#pragma unroll
      for (int q = 0; q < 4; q++) out[(2 + q) * nm + i] = su[q][i];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ k_dceig
// Eigenvalues of the tridiagonal T and the projections of the transformed
// coil and rho0 on its eigenvectors, by divide and conquer (Cuppen's method
// with the eigenvector formula of Gu & Eisenstat): T is torn into leaves of
// at most kDcLeaf rows by rank-one tears, T = diag(T1, T2) + |b| w w^T
// (w = e_m + sign(b) e_{m+1}, b the coupling at the tear); each leaf is
// solved by implicit QR with Wilkinson shifts (one thread per leaf); the
// leaves are merged level by level, every merge solving
//   D + rho z z^T,  rho = 2|b|,  z = [last row of Q1, sign(b) first row of Q2]/sqrt(2)
// through the secular equation 1 + rho sum z_i^2/(d_i - lambda) = 0 (one
// thread per root). Eigenvectors are never stored: each subproblem carries
// six vectors over its eigen-index, Q^T applied to conj(coil) and rho0 (real
// and imaginary parts) and its first and last eigenvector rows (needed by the
// parent merge). One 256-thread workgroup per orientation, everything in LDS.
constexpr int kDcThreads = 256;
constexpr int kDcLeaf = 6;
constexpr int kDcMaxLevels = 7;  // up to 128 leaves (kTriGNP)
constexpr int kDcMaxIter = 100;  // secular-equation iterations per root (typically 3-8)
constexpr int kDcPass1 = 3;      // evaluations per root before the open roots are packed (see M3a; 2 and 4-5 measured slower)

// Sized for tridiagonal matrices of dimension at most NM (kTriNP, kTriWNP or kTriGNP), solved
// by kThr threads, one per root of a merge.
template <int NM>
struct DcSharedT {
  static constexpr int kNM = NM;
  static constexpr int kThr = (NM <= kDcThreads) ? kDcThreads : 2 * kDcThreads;
  static constexpr int kLev = (NM <= kTriNP) ? 5 : ((NM <= kTriWNP) ? 6 : kDcMaxLevels);  // leaves of <= kDcLeaf rows
  double d[NM], e[NM];
  double X[6][NM];  // tracked vectors over the eigen-index of each segment
  double Y[6][NM];  // merge input, sorted
  double D[NM], z[NM];
  double Dc[NM], Zc[NM], Z2c[NM], zh[NM], tau[NM];
  double lorg[NM];  // Dc[org] of each root: lambda = lorg + tau
  int org[NM], keep[NM], defl[NM];
  int seg[NM];  // merge index of each element at the current level (set by M1)
  int bnd[kLev + 1][(1 << kLev) + 1];  // segment boundaries per level
  double beta[(1 << kLev)];            // rho = 2|coupling| of the merges of a level
  double zsum[(1 << kLev)];            // sum of the kept z^2 of each merge
  int K[(1 << kLev)];
  int its[NM];
  unsigned long long pend[kThr / 64];  // roots the first secular pass left open, by thread
  int nlev;
  int fail;
  int max_iter;  // secular-equation evaluations allowed per root in the second pass
};
using DcShared = DcSharedT<kTriNP>;
static_assert(kTriGNP <= 2 * kDcThreads && kDcLeaf << kDcMaxLevels >= kTriGNP, "eigensolver sizes");

// 1/x: on the device v_rcp_f64 refined by two Newton steps (full double
// precision from the hardware estimate), on the host a division
__host__ __device__ inline double dc_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
#else
  return 1.0 / x;
#endif
}

// 1/x to within 2^-48.8, about 18 ulp (one Newton step on v_rcp_f64's 2^-24.4): for the
// term sums of the secular iteration and of its starting point (the sign of f at the
// midpoint), where it is below the rounding bound of the K-term sums from K = 19 on; on the
// 8-spin workload the FIDs then differ from the oracle by 9.0e-13 instead of 8.3e-13 (batch
// and dense-generator cases unchanged). Everything else (the model roots, z recomputation,
// eigenvector formula) uses dc_rcp.
__host__ __device__ inline double dc_rcp_iter(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
#else
  return 1.0 / x;
#endif
}

// The kernel is a sequence of phases separated by workgroup barriers; within
// a phase every thread works on its own data. Each phase is a
// __host__ __device__ function of the thread id, so that a CPU build can run
// the same code phase by phase (a loop over thread ids per phase).

// implicit QR on the segment [a, b) of (d, e) with the six tracked vectors
// (Golub & Van Loan, algorithm 8.3.3); deflation at |e_i| <= eps*|T_seg|
template <class SH>
__host__ __device__ inline int dc_leaf_qr(SH& S, int a, int b) {
  const int n = b - a;
  double anorm = 0.0;
  for (int i = a; i < b; i++) {
    double r = fabs(S.d[i]);
    if (i + 1 < b) r += fabs(S.e[i]);
    if (i > a) r += fabs(S.e[i - 1]);
    anorm = fmax(anorm, r);
  }
  const double tol = 2.220446049250313e-16 * anorm;
  int hi = b - 1, sweeps = 0;
  while (hi > a) {
    if (fabs(S.e[hi - 1]) <= tol) {
      S.e[hi - 1] = 0.0;
      hi--;
      continue;
    }
    int lo = hi - 1;
    while (lo > a && fabs(S.e[lo - 1]) > tol) lo--;
    if (++sweeps > 40 * n) return -1;
    const double dd = 0.5 * (S.d[hi - 1] - S.d[hi]), bb = S.e[hi - 1];
    const double h = sqrt(dd * dd + bb * bb);
    const double mu = S.d[hi] - bb * bb / (dd + (dd >= 0.0 ? h : -h));
    double x = S.d[lo] - mu, zz = S.e[lo];
    for (int k = lo; k < hi; k++) {
      const double r = sqrt(x * x + zz * zz);
      double c = 1.0, s = 0.0;
      if (r > 0.0) {
        c = x / r;
        s = zz / r;
      }
      if (k > lo) S.e[k - 1] = r;
      const double dk = S.d[k], dk1 = S.d[k + 1], ek = S.e[k];
      const double cc = c * c, ss = s * s, cs = c * s;
      S.d[k] = cc * dk + 2.0 * cs * ek + ss * dk1;
      S.d[k + 1] = ss * dk - 2.0 * cs * ek + cc * dk1;
      S.e[k] = cs * (dk1 - dk) + (cc - ss) * ek;
      if (k < hi - 1) {
        x = S.e[k];
        zz = s * S.e[k + 1];
        S.e[k + 1] = c * S.e[k + 1];
      }
      for (int q = 0; q < 6; q++) {
        const double u0 = S.X[q][k], u1 = S.X[q][k + 1];
        S.X[q][k] = c * u0 + s * u1;
        S.X[q][k + 1] = c * u1 - s * u0;
      }
    }
  }
  return sweeps;
}

// insertion sort of the segment [a, b) by eigenvalue, carrying the vectors
template <class SH>
__host__ __device__ inline void dc_sort_segment(SH& S, int a, int b) {
  for (int i = a + 1; i < b; i++) {
    const double v = S.d[i];
    double x[6];
    for (int q = 0; q < 6; q++) x[q] = S.X[q][i];
    int j = i - 1;
    while (j >= a && S.d[j] > v) {
      S.d[j + 1] = S.d[j];
      for (int q = 0; q < 6; q++) S.X[q][j + 1] = S.X[q][j];
      j--;
    }
    S.d[j + 1] = v;
    for (int q = 0; q < 6; q++) S.X[q][j + 1] = x[q];
  }
}

// Root k of 1/rho + sum_{i<K} Z2_i/(Dc_i - lambda) = 0 (Dc ascending, Z2 = z^2 > 0,
// zs = sum Z2) as lambda = Dc[org] + tau, org the nearer pole. Start: the two-pole model
// through poles kk, kk+1 with the other terms frozen at the middle of the interval (the
// interval (Dc[K-1], Dc[K-1] + rho zs) for the outermost root), as LAPACK dlaed4 starts.
// Steps: the rational models of R.-C. Li ("Solving secular equations stably and
// efficiently", 1993) through the same two poles -- fixed weight (the origin pole's own
// weight, the other terms' slope given to the far pole) and middle way (the slopes of the
// two sides given to the two poles) -- switched when |f| does not fall tenfold with an
// unchanged sign; for the outermost root the model's root beyond the last pole. Both models
// match f and f' at the current point, so they converge at least quadratically: a step
// below 1e-7 |tau| is taken and ends the iteration (1e-9 costs 2.5% more time for FIDs equal
// to three digits; 1e-6 raised the dense-generator error from 1.1e-12 to 6.8e-12). The
// bracket guards every step (a
// bisection replaces a step that leaves it). dc_secular_start and dc_secular_run below
// split this into the start and a resumable iteration.

// The state of one root's iteration, so that it can stop after a given number of
// evaluations of f and resume (in another lane) later.
struct SecState {
  double tau, tlo, thi, fprev;
  int org, it, fixed_weight;
};

// The start of root k (org, tau, bracket); returns true if that is already the
// root (K == 1: lambda = d + rho z^2).
__host__ __device__ __attribute__((always_inline)) inline bool dc_secular_start(const double* Dc, const double* Z2,
                                                                                 int K, int k, double rho, double zs,
                                                                                 SecState& st) {
  const double rinv = 1.0 / rho;
  st.it = 1;
  st.fixed_weight = 1;
  st.fprev = 0.0;
  if (K == 1) {
    st.org = 0;
    st.tau = rho * Z2[0];
    st.tlo = 0.0;
    st.thi = st.tau;
    return true;
  }
  const bool outer = (k == K - 1);
  const int kk = outer ? K - 2 : k;  // the model's poles: kk and kk + 1
  const double gap = Dc[kk + 1] - Dc[kk];
  const int base = outer ? K - 1 : k;
  const double x0 = outer ? 0.5 * rho * zs : 0.5 * gap;
  const double dbase = Dc[base];
  double f = rinv;
#pragma unroll 4
  for (int i = 0; i < K; i++) f += Z2[i] * dc_rcp_iter((Dc[i] - dbase) - x0);
  const double s2 = Z2[kk] + Z2[kk + 1];
  // the bracket and the origin as values (no branch assigns them, so they stay in registers)
  const bool left = outer ? true : (f >= 0.0);  // interior: root in (Dc[k], midpoint]
  const int org = outer ? K - 1 : (left ? k : k + 1);
  const double tlo = outer ? ((f > 0.0) ? 0.0 : x0) : (left ? 0.0 : -x0);
  const double thi = outer ? ((f > 0.0) ? x0 : 2.0 * x0) : (left ? x0 : 0.0);
  double tau;
  if (!outer) {
    // the other terms at the midpoint: f less Z2_kk/(-x0) and Z2_kk+1/x0
    const double c = f + (Z2[kk] - Z2[kk + 1]) * dc_rcp(x0);
    if (left) {  // c t^2 - bq t + aq = 0, t = lambda - Dc[k]
      const double bq = c * gap + s2, aq = Z2[kk] * gap;
      const double r = sqrt(fabs(bq * bq - 4.0 * aq * c));
      tau = (bq > 0.0) ? 2.0 * aq / (bq + r) : (bq - r) / (2.0 * c);
    } else {  // root in (midpoint, Dc[k+1]): c t^2 - aq t - bq = 0, t = lambda - Dc[k+1]
      const double aq = s2 - c * gap, bq = Z2[kk + 1] * gap;
      const double r = sqrt(fabs(aq * aq + 4.0 * c * bq));
      tau = (aq > 0.0) ? -2.0 * bq / (aq + r) : (aq - r) / (2.0 * c);
    }
  } else {  // c t^2 - aq t - bq = 0 for t = lambda - Dc[K-1] > 0
    const double c = f + Z2[kk] * dc_rcp(gap + x0) + Z2[kk + 1] * dc_rcp(x0);
    const double aq = s2 - c * gap, bq = Z2[kk + 1] * gap;
    const double r = sqrt(fabs(aq * aq + 4.0 * c * bq));
    tau = (aq >= 0.0) ? (aq + r) / (2.0 * c) : 2.0 * bq / (r - aq);
  }
  if (!(tau > tlo && tau < thi)) tau = 0.5 * (tlo + thi);
  st.org = org;
  st.tau = tau;
  st.tlo = tlo;
  st.thi = thi;
  return false;
}

// Evaluations of f and model steps from the state st until root k has converged
// (returns true; st.tau final, st.it the evaluations used) or st.it exceeds
// limit (returns false; st holds everything needed to resume).
__host__ __device__ __attribute__((always_inline)) inline bool dc_secular_run(const double* Dc, const double* Z2, int K,
                                                                               int k, double rho, SecState& st,
                                                                               int limit) {
  const double eps = 2.220446049250313e-16;
  const double rinv = 1.0 / rho;
  const bool outer = (k == K - 1);
  const int kk = outer ? K - 2 : k;  // the model's poles: kk and kk + 1
  const double gap = Dc[kk + 1] - Dc[kk];
  const int org = st.org;
  double tau = st.tau, tlo = st.tlo, thi = st.thi, fprev = st.fprev;
  bool fixed_weight = st.fixed_weight != 0;
  const double dorg = Dc[org];
  const double e1 = Dc[kk] - dorg, e2 = Dc[kk + 1] - dorg;  // the model's poles from the origin
  int it = st.it;
  bool done = false;
  for (; it <= limit; it++) {
    // f, sum |w|, f' = sum w/delta and sum w/|delta|: the poles below lambda (psi side,
    // delta < 0) enter the last with a minus sign, so the two slopes are the half-difference
    // and half-sum of the derivative sums (two FMAs instead of a split by sign)
    double fs = 0.0, wabs = 0.0, dw = 0.0, sq = 0.0;
#pragma unroll 4
    for (int i = 0; i < K - 1; i++) {
      const double rd = dc_rcp_iter((Dc[i] - dorg) - tau);
      const double w = Z2[i] * rd;
      fs += w;
      wabs += fabs(w);
      dw = fma(w, rd, dw);
      sq = fma(w, fabs(rd), sq);
    }
    // the last pole: the far pole of the outermost root's model, else a term as the rest
    const double rl = dc_rcp((Dc[K - 1] - dorg) - tau);
    const double wl = Z2[K - 1] * rl;
    fs += wl;
    wabs += fabs(wl);
    double psip, phip;
    if (outer) {
      psip = dw;  // every other pole lies below lambda
      phip = wl * rl;
      dw += phip;
    } else {
      dw = fma(wl, rl, dw);
      sq = fma(wl, fabs(rl), sq);
      psip = 0.5 * (dw - sq);
      phip = 0.5 * (dw + sq);
    }
    const double fv = rinv + fs;
    if (fv > 0.0)
      thi = tau;
    else
      tlo = tau;
    if (fabs(fv) <= 8.0 * eps * (rinv + wabs) || thi - tlo <= 4.0 * eps * fmax(fabs(tlo), fabs(thi))) {
      done = true;
      break;
    }
    const double d1 = e1 - tau, d2 = e2 - tau;
    if (fv * fprev > 0.0 && fabs(fv) > 0.1 * fabs(fprev)) fixed_weight = !fixed_weight;
    fprev = fv;
    double c;
    if (!outer && fixed_weight) {
      const bool lo = (org == k);
      const double dn = lo ? d1 : d2, dfar = lo ? d2 : d1;
      const double r1 = dc_rcp(dn);
      const double zo = lo ? Z2[kk] : Z2[kk + 1];
      c = fv - dfar * dw - (lo ? -gap : gap) * (zo * r1 * r1);
    } else {
      c = fv - psip * d1 - phip * d2;
    }
    const double B = (d1 + d2) * fv - d1 * d2 * dw, C = d1 * d2 * fv;
    const double disc = sqrt(fabs(B * B - 4.0 * c * C));
    double eta;
    if (!outer)
      eta = (B >= 0.0) ? 2.0 * C / (B + disc) : (B - disc) / (2.0 * c);
    else
      eta = (B >= 0.0) ? (B + disc) / (2.0 * fabs(c)) : 2.0 * C / (B - disc);
    const double nt = tau + eta;
    if (!(nt > tlo && nt < thi)) {
      tau = 0.5 * (tlo + thi);
      continue;
    }
    const bool small = fabs(eta) <= 1e-7 * fabs(tau);
    tau = nt;
    if (small) {
      done = true;
      break;
    }
  }
  st.tau = tau;
  st.tlo = tlo;
  st.thi = thi;
  st.fprev = fprev;
  st.fixed_weight = fixed_weight ? 1 : 0;
  st.it = it;
  return done;
}



// merge index of element i at level l (bounded by the level's segment count)
template <class SH>
__host__ __device__ inline int dc_segment_of(const SH& S, int l, int i) {
  const int last = (1 << l) - 1;
  int s2 = 0;
  while (s2 < last && S.bnd[l][s2 + 1] <= i) s2++;
  return s2;
}

// phase 0: load T and the four vectors (every thread)
template <class SH>
__host__ __device__ inline void dc_phase_load(SH& S, const double* gsrc, int nm, int tid, int nthreads) {
  for (int i = tid; i < SH::kNM; i += nthreads) {
    S.d[i] = gsrc[i];
    S.e[i] = gsrc[nm + i];
    for (int q = 0; q < 4; q++) S.X[q][i] = gsrc[(2 + q) * nm + i];
    S.X[4][i] = 0.0;
    S.X[5][i] = 0.0;
    S.its[i] = 0;
  }
}

// phase 1 (one thread): the segment tree and the rank-one tears
template <class SH>
__host__ __device__ inline void dc_phase_plan(SH& S, int n, int max_iter = kDcMaxIter) {
  int L = 0;
  while (L < SH::kLev && ((n + (1 << L) - 1) >> L) > kDcLeaf) L++;
  S.nlev = L;
  S.fail = 0;
  S.max_iter = max_iter;
  S.bnd[0][0] = 0;
  S.bnd[0][1] = n;
  for (int l = 0; l < L; l++) {
    for (int s2 = 0; s2 < (1 << l); s2++) {
      const int a = S.bnd[l][s2], b = S.bnd[l][s2 + 1];
      S.bnd[l + 1][2 * s2] = a;
      S.bnd[l + 1][2 * s2 + 1] = a + (b - a) / 2;
    }
    S.bnd[l + 1][1 << (l + 1)] = n;
  }
  for (int l = 0; l < L; l++)
    for (int s2 = 0; s2 < (1 << l); s2++) {
      const int m = S.bnd[l + 1][2 * s2 + 1];
      if (m <= S.bnd[l][s2] || m >= S.bnd[l][s2 + 1]) continue;  // empty child
      const double b = fabs(S.e[m - 1]);
      S.d[m - 1] -= b;
      S.d[m] -= b;
    }
}

// phase 2: one thread per leaf
template <class SH>
__host__ __device__ inline void dc_phase_leaf(SH& S, int tid) {
  const int L = S.nlev;
  if (tid >= (1 << L)) return;
  const int a = S.bnd[L][tid], b = S.bnd[L][tid + 1];
  if (b <= a) return;
  S.X[4][a] = 1.0;
  S.X[5][b - 1] = 1.0;
  if (dc_leaf_qr(S, a, b) < 0) S.fail = 1;
  dc_sort_segment(S, a, b);
}

// M1: merge-path of the two sorted children into Y, D, z (one thread per element)
template <class SH>
__host__ __device__ inline void dc_phase_m1(SH& S, int l, int n, int i) {
  if (i >= n) return;
  const int s2 = dc_segment_of(S, l, i);
  S.seg[i] = s2;
  const int a = S.bnd[l][s2], b = S.bnd[l][s2 + 1], m = S.bnd[l + 1][2 * s2 + 1];
  const double bcoup = (m > a && m < b) ? S.e[m - 1] : 0.0;
  const double sg = (bcoup >= 0.0) ? 1.0 : -1.0;
  const bool left = (i < m);
  const double v = S.d[i];
  // rank in the other child: a left element counts right < v, a right one left <= v
  int lo2 = left ? m : a, hi2 = left ? b : m;
  while (lo2 < hi2) {
    const int md = (lo2 + hi2) >> 1;
    const bool before = left ? (S.d[md] < v) : (S.d[md] <= v);
    if (before)
      lo2 = md + 1;
    else
      hi2 = md;
  }
  const int pos = a + (i - (left ? a : m)) + (lo2 - (left ? m : a));
  S.D[pos] = v;
  S.z[pos] = (left ? S.X[5][i] : sg * S.X[4][i]) * 0.70710678118654752440;
  for (int q = 0; q < 4; q++) S.Y[q][pos] = S.X[q][i];
  S.Y[4][pos] = left ? S.X[4][i] : 0.0;
  S.Y[5][pos] = left ? 0.0 : S.X[5][i];
}

// M2: deflation scan (one thread per merge): small z_i, and pairs of close
// poles rotated so that one z component vanishes; the kept poles are compacted
// into Dc, Zc, Z2c as they are found (a later rotation only removes the last kept
// one, whose slot the rotated pole takes)
template <class SH>
__host__ __device__ inline void dc_phase_m2(SH& S, int l, int s2) {
  if (s2 >= (1 << l)) return;
  if (s2 == 0)
    for (int w = 0; w < SH::kThr / 64; w++) S.pend[w] = 0;
  const int a = S.bnd[l][s2], b = S.bnd[l][s2 + 1], m = S.bnd[l + 1][2 * s2 + 1];
  const double bcoup = (m > a && m < b) ? S.e[m - 1] : 0.0;
  const double rho = 2.0 * fabs(bcoup);
  S.beta[s2] = rho;
  // D is ascending on [a, b): the largest magnitude is at one of the ends
  const double dmax = (b > a) ? fmax(fabs(S.D[a]), fabs(S.D[b - 1])) : 0.0;
  const double tol = 8.0 * 2.220446049250313e-16 * fmax(dmax, rho);
  int K = 0, nd = 0, prev = -1;
  double zs = 0.0;  // sum of the kept z^2 (a rotation keeps the sum of the pair's squares)
  // the last kept element's z and D as they stand are carried in registers, and the next
  // element's are read one iteration ahead (an iteration writes only elements <= i)
  double zp = 0.0, dp = 0.0;
  double zn = (a < b) ? S.z[a] : 0.0, dn = (a < b) ? S.D[a] : 0.0;
  for (int i = a; i < b; i++) {
    const double zi = zn, di = dn;
    if (i + 1 < b) {
      zn = S.z[i + 1];
      dn = S.D[i + 1];
    }
    if (rho == 0.0 || rho * fabs(zi) <= tol) {
      S.defl[a + nd++] = i;
      continue;
    }
    zs = fma(zi, zi, zs);
    double zk = zi, dk = di;
    if (prev >= 0) {
      const double t = di - dp;
      const double zz = fma(zp, zp, zi * zi);
      // |t c s| <= tol for the rotation c = zi/|z|, s = -zp/|z|, tested without the root
      if (fabs(t * zi * zp) <= tol * zz) {
        const double tz = sqrt(zz);
        const double c = zi / tz, s = -zp / tz;
        zk = tz;
        S.z[i] = tz;
        S.z[prev] = 0.0;
        S.D[prev] = dp * c * c + di * s * s;
        dk = dp * s * s + di * c * c;
        S.D[i] = dk;
        for (int q = 0; q < 6; q++) {
          const double u0 = S.Y[q][prev], u1 = S.Y[q][i];
          S.Y[q][prev] = c * u0 + s * u1;
          S.Y[q][i] = c * u1 - s * u0;
        }
        S.defl[a + nd++] = prev;
        K--;
      }
    }
    S.keep[a + K] = i;
    S.Dc[a + K] = dk;
    S.Zc[a + K] = zk;
    S.Z2c[a + K] = zk * zk;
    K++;
    prev = i;
    zp = zk;
    dp = dk;
  }
  S.K[s2] = K;
  S.zsum[s2] = zs;
}

// the position of the r-th (from 0) set bit of m
__host__ __device__ inline int dc_select_bit(unsigned long long m, int r) {
  int pos = 0;
  for (int width = 32; width >= 1; width >>= 1) {
    const unsigned long long low = m & ((1ull << width) - 1);
    const int c = __builtin_popcountll(low);
    if (r >= c) {
      r -= c;
      m >>= width;
      pos += width;
    } else {
      m = low;
    }
  }
  return pos;
}

template <class SH>
__host__ __device__ inline void dc_root_store(SH& S, int i, int a, int k, const SecState& st) {
  if (st.it > S.its[i]) S.its[i] = st.it;
  S.org[a + k] = st.org;
  S.tau[a + k] = st.tau;
  S.lorg[a + k] = S.Dc[a + st.org];
}

// M3a: secular roots, first pass (one thread per (merge, root)): at most kDcPass1
// evaluations of f per root. A root still open leaves its state in X (unused
// between M1 and M5) at its thread index and its bit in pend, so that M3b can
// finish the open roots packed into as few waves as they fill: a wave's
// iterations cost its slowest lane, and most roots converge in 2-4.
template <class SH>
__host__ __device__ inline void dc_phase_m3a(SH& S, int l, int n, int i) {
  if (i >= n) return;
  const int s2 = S.seg[i];
  const int a = S.bnd[l][s2], k = i - a, K = S.K[s2];
  if (k >= K) return;
  SecState st;
  bool done = dc_secular_start(S.Dc + a, S.Z2c + a, K, k, S.beta[s2], S.zsum[s2], st);
  if (!done) done = dc_secular_run(S.Dc + a, S.Z2c + a, K, k, S.beta[s2], st, kDcPass1);
  if (done) {
    dc_root_store(S, i, a, k, st);
    return;
  }
  S.X[0][i] = st.tau;
  S.X[1][i] = st.tlo;
  S.X[2][i] = st.thi;
  S.X[3][i] = st.fprev;
  S.X[4][i] = (double)st.org;
  S.X[5][i] = (double)(2 * st.it + st.fixed_weight);
#if defined(__HIP_DEVICE_COMPILE__)
  atomicOr(&S.pend[i >> 6], 1ull << (i & 63));
#else
  S.pend[i >> 6] |= 1ull << (i & 63);
#endif
}

// M3b: thread u finishes the u-th open root (in thread order)
template <class SH>
__host__ __device__ inline void dc_phase_m3b(SH& S, int l, int u) {
  int w = 0, base = 0;
  for (; w < SH::kThr / 64; w++) {
    const int c = __builtin_popcountll(S.pend[w]);
    if (u < base + c) break;
    base += c;
  }
  if (w == SH::kThr / 64) return;
  const int i = 64 * w + dc_select_bit(S.pend[w], u - base);
  const int s2 = S.seg[i];
  const int a = S.bnd[l][s2], k = i - a, K = S.K[s2];
  SecState st;
  st.tau = S.X[0][i];
  st.tlo = S.X[1][i];
  st.thi = S.X[2][i];
  st.fprev = S.X[3][i];
  st.org = (int)S.X[4][i];
  const int code = (int)S.X[5][i];
  st.it = code >> 1;
  st.fixed_weight = code & 1;
  if (!dc_secular_run(S.Dc + a, S.Z2c + a, K, k, S.beta[s2], st, S.max_iter)) S.fail = 1;
  dc_root_store(S, i, a, k, st);
}

// M4: z recomputed from the computed roots (Gu & Eisenstat), one thread per (merge, pole):
// z_i^2 = (lambda_i - d_i)/rho prod_{k != i} (lambda_k - d_i)/(d_k - d_i), the products
// taken four factors at a time with one reciprocal per group
template <class SH>
__host__ __device__ inline void dc_phase_m4(SH& S, int l, int n, int t) {
  if (t >= n) return;
  const int s2 = S.seg[t];
  const int a = S.bnd[l][s2], i = t - a, K = S.K[s2];
  if (i >= K) return;
  const double di = S.Dc[a + i];
  double w = 1.0, num = 1.0, den = 1.0;
#pragma unroll 4
  for (int k = 0; k < K; k++) {
    num *= (S.lorg[a + k] - di) + S.tau[a + k];  // lambda_k - d_i
    den *= (k != i) ? S.Dc[a + k] - di : 1.0;    // d_k - d_i
    if ((k & 3) == 3) {
      w *= num * dc_rcp(den);
      num = 1.0;
      den = 1.0;
    }
  }
  w *= num * dc_rcp(den);
  const double zz = sqrt(fmax(w, 0.0) / S.beta[s2]);
  S.zh[a + i] = (S.Zc[a + i] >= 0.0) ? zz : -zz;
}

// M5: eigenpairs of each merge into d, X: roots in slots 0..K-1, deflated pairs after
template <class SH>
__host__ __device__ inline void dc_phase_m5(SH& S, int l, int n, int t) {
  if (t >= n) return;
  const int s2 = S.seg[t];
  const int a = S.bnd[l][s2], slot = t - a, K = S.K[s2];
  double out[6];
  double lam;
  if (slot < K) {
    const double tu = S.tau[a + slot], dorg = S.lorg[a + slot];
    double nrm = 0.0;
    for (int q = 0; q < 6; q++) out[q] = 0.0;
    if (l > 0) {
      for (int i = 0; i < K; i++) {
        const double u = S.zh[a + i] * dc_rcp((S.Dc[a + i] - dorg) - tu);
        nrm = fma(u, u, nrm);
        const int src = S.keep[a + i];
        for (int q = 0; q < 6; q++) out[q] = fma(u, S.Y[q][src], out[q]);
      }
    } else {  // the top merge has no parent: its first and last eigenvector rows are not needed
      for (int i = 0; i < K; i++) {
        const double u = S.zh[a + i] * dc_rcp((S.Dc[a + i] - dorg) - tu);
        nrm = fma(u, u, nrm);
        const int src = S.keep[a + i];
        for (int q = 0; q < 4; q++) out[q] = fma(u, S.Y[q][src], out[q]);
      }
    }
    const double inv = 1.0 / sqrt(nrm);
    for (int q = 0; q < 6; q++) out[q] *= inv;
    lam = dorg + tu;
  } else {
    const int src = S.defl[a + slot - K];
    for (int q = 0; q < 6; q++) out[q] = S.Y[q][src];
    lam = S.D[src];
  }
  S.d[t] = lam;
  for (int q = 0; q < 6; q++) S.X[q][t] = out[q];
}

// M6a/M6b: rank sort of each merged segment (the rank into stage_*, then the scatter)
// (the ranks and copies are staged in org, D and Y, which the merge no longer needs)
template <class SH>
__host__ __device__ inline void dc_phase_m6a(SH& S, int l, int n, int t) {
  if (t >= SH::kNM) return;
  S.org[t] = -1;
  if (t >= n) return;
  const int s2 = S.seg[t];
  const int a = S.bnd[l][s2], b = S.bnd[l][s2 + 1];
  const double lam = S.d[t];
  int r = 0;
  for (int j = a; j < b; j++) {
    const double v = S.d[j];
    r += (v < lam || (v == lam && j < t)) ? 1 : 0;
  }
  S.org[t] = a + r;
  S.D[t] = lam;
  for (int q = 0; q < 6; q++) S.Y[q][t] = S.X[q][t];
}

template <class SH>
__host__ __device__ inline void dc_phase_m6b(SH& S, int t) {
  if (t >= SH::kNM) return;
  const int pos = S.org[t];
  if (pos < 0) return;
  S.d[pos] = S.D[t];
  for (int q = 0; q < 6; q++) S.X[q][pos] = S.Y[q][t];
}

// output: eigenvalues and g_k = a_k * b_k (a = transformed conj(coil), b = transformed rho0)
template <class SH>
__host__ __device__ inline void dc_phase_store(const SH& S, double* gsrc, int nm, int n, int tid, int nthreads) {
  for (int k = tid; k < n; k += nthreads) {
    const double ar = S.X[0][k], ai = S.X[1][k], br = S.X[2][k], bi = S.X[3][k];
    gsrc[k] = S.d[k];
    gsrc[2 * nm + k] = ar * br - ai * bi;
    gsrc[3 * nm + k] = ar * bi + ai * br;
  }
}

template <int NM>
__global__ void __launch_bounds__(DcSharedT<NM>::kThr) k_dceig(SpecParams p) {
  __shared__ DcSharedT<NM> S;
  constexpr int kT = DcSharedT<NM>::kThr;
  const int tid = threadIdx.x;
  const int64_t it = blockIdx.x;
  if (it >= p.n_items) return;
  const int n = p.sys[p.items[it].sys].n;
  const int nm = p.spec_nm;
  double* gsrc = p.spec + (size_t)it * 6 * nm;
  dc_phase_load(S, gsrc, nm, tid, kT);
  __syncthreads();
  if (tid == 0) dc_phase_plan(S, n, p.dc_max_iter > 0 ? p.dc_max_iter : kDcMaxIter);
  __syncthreads();
  // the leaves and the deflation scans are serial on a few lanes of wave 0 while the
  // rest of the workgroup waits: their wave runs ahead of other workgroups' waves
  if (tid < 64) __builtin_amdgcn_s_setprio(3);
  dc_phase_leaf(S, tid);
  if (tid < 64) __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  for (int l = S.nlev - 1; l >= 0; l--) {
    dc_phase_m1(S, l, n, tid);
    __syncthreads();
    if (tid < 64) __builtin_amdgcn_s_setprio(3);
    dc_phase_m2(S, l, tid);
    if (tid < 64) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    dc_phase_m3a(S, l, n, tid);
    __syncthreads();
    dc_phase_m3b(S, l, tid);
    __syncthreads();
    dc_phase_m4(S, l, n, tid);
    __syncthreads();
    dc_phase_m5(S, l, n, tid);
    __syncthreads();
    if (l > 0) {
      dc_phase_m6a(S, l, n, tid);
      __syncthreads();
      dc_phase_m6b(S, tid);
      __syncthreads();
    }
  }
  dc_phase_store(S, gsrc, nm, n, tid, kT);
  __syncthreads();
  if (tid == 0) {
    int mx = 0;
    for (int i = 0; i < NM; i++) mx = (S.its[i] > mx) ? S.its[i] : mx;
    p.status[it] = S.fail ? -1 : mx;
  }
}

// ------------------------------------------------------------ k_synth
// f(64 (i0 + i) + j) = sum_k U[i][k] W[k][j] for one row tile i < 64 of
// 64-point rows, U[i][k] = g_k e^{-i lambda_k 64 (i0 + i)}, W[k][j] =
// e^{-i lambda_k j}: a complex 64 x 64 x np product on the matrix cores, its
// operands generated panel by panel straight into the LDS staging buffers.
// Each eigenvalue's phase steps e^{-i lambda m} for m = 1, 8, 64, 512 come from
// one sincos each; an entry is its group's start (a binary power of the 8- or
// 512-step, at most four rounded products) advanced by the 1- or 64-step (at
// most seven more), so every phase factor carries at most ~12 roundings.
template <int NM>
struct SynthLds {
  double s1r[NM], s1i[NM], s8r[NM], s8i[NM];
  double s64r[NM], s64i[NM], s512r[NM], s512i[NM];
  double gr[NM], gi[NM];  // g_k e^{-i lambda_k 64 i0} for the current row tile
  GemmLds<double> g;
  double red[kThreads / 64][2];  // per-wave parts of a tail point
  int sitem;
};

__device__ __forceinline__ void cmul(double& ar, double& ai, double br, double bi) {
  const double t = ar * br - ai * bi;
  ai = ar * bi + ai * br;
  ar = t;
}

template <int NM>
__global__ void __launch_bounds__(kThreads, 3) k_synth(SpecParams p) {
  typedef Mma<double>::v4 v4;
  __shared__ SynthLds<NM> S;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;  // the wave's 32 x 32 quarter of the 64 x 64 tile
  const int64_t npts = p.nsteps + 1;
  const int64_t nfull = npts / kSynthB;  // rows of 64 points
  for (;;) {
    if (tid == 0) S.sitem = (int)atomicAdd(&p.queue[1], 1u);
    __syncthreads();
    const int64_t it = S.sitem;
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sdesc = p.sys[item.sys];
    const int n = sdesc.n, np = sdesc.np;
    const double* lam = p.spec + (size_t)it * 6 * p.spec_nm;
    const double* gre = lam + 2 * p.spec_nm;
    const double* gim = lam + 3 * p.spec_nm;
    double* out = p.per_or + 2 * (size_t)item.orient * npts;
    for (int k = tid; k < np; k += kThreads) {
      double c1 = 1.0, n1 = 0.0, c8 = 1.0, n8 = 0.0, c64 = 1.0, n64 = 0.0, c512 = 1.0, n512 = 0.0;
      if (k < n) {
        double sn;
        sincos(lam[k], &sn, &c1);
        n1 = -sn;
        sincos(lam[k] * 8.0, &sn, &c8);
        n8 = -sn;
        sincos(lam[k] * 64.0, &sn, &c64);
        n64 = -sn;
        sincos(lam[k] * 512.0, &sn, &c512);
        n512 = -sn;
      }
      S.s1r[k] = c1;
      S.s1i[k] = n1;
      S.s8r[k] = c8;
      S.s8i[k] = n8;
      S.s64r[k] = c64;
      S.s64i[k] = n64;
      S.s512r[k] = c512;
      S.s512i[k] = n512;
    }
    for (int64_t i0 = 0; i0 < nfull; i0 += kSynthB) {
      const int rows = (int)((nfull - i0 < kSynthB) ? (nfull - i0) : kSynthB);
      for (int k = tid; k < np; k += kThreads) {
        double br = 0.0, bi = 0.0;
        if (k < n) {
          double sn = 0.0, cs = 1.0;
          if (i0 > 0) sincos(lam[k] * (double)(kSynthB * i0), &sn, &cs);
          br = gre[k] * cs + gim[k] * sn;
          bi = gim[k] * cs - gre[k] * sn;
        }
        S.gr[k] = br;
        S.gi[k] = bi;
      }
      v4 cr[2][2], ci[2][2];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          cr[a][b] = v4{0, 0, 0, 0};
          ci[a][b] = v4{0, 0, 0, 0};
        }
      const bool active = (32 * wr < rows);
      for (int k0 = 0; k0 < np; k0 += 16) {
        __syncthreads();  // the step tables are written / the previous panel is consumed
        {
          // waves 0-1: W[k0 + kk][8a .. 8a+7]; waves 2-3: U[8a .. 8a+7][k0 + kk]
          const bool wpanel = (tid < 128);
          const int t7 = tid & 127, kk = t7 >> 3, a = t7 & 7, k = k0 + kk;
          double qr, qi, sr, si, br, bi;
          if (wpanel) {
            qr = S.s8r[k];
            qi = S.s8i[k];
            sr = S.s1r[k];
            si = S.s1i[k];
            br = 1.0;
            bi = 0.0;
          } else {
            qr = S.s512r[k];
            qi = S.s512i[k];
            sr = S.s64r[k];
            si = S.s64i[k];
            br = S.gr[k];
            bi = S.gi[k];
          }
          // the group start: base * q^a
#pragma unroll
          for (int m = 0; m < 3; m++) {
            if ((a >> m) & 1) cmul(br, bi, qr, qi);
            if (m < 2) cmul(qr, qi, qr, qi);
          }
#pragma unroll
          for (int b = 0; b < 8; b++) {
            const int e = 8 * a + b;
            if (wpanel) {
              S.g.br[kk][e] = br;
              S.g.bi[kk][e] = bi;
            } else {
              const bool ok = (e < rows);
              S.g.ar[kk][e] = ok ? br : 0.0;
              S.g.ai[kk][e] = ok ? bi : 0.0;
            }
            cmul(br, bi, sr, si);
          }
        }
        __syncthreads();
        if (active) {
#pragma unroll
          for (int kk = 0; kk < 16; kk += 4) {
            const int kr = kk + (lane >> 4);
            double xr[2], xi[2], nxi[2], yr[2], yi[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
              xr[t] = S.g.ar[kr][32 * wr + 16 * t + (lane & 15)];
              xi[t] = S.g.ai[kr][32 * wr + 16 * t + (lane & 15)];
              nxi[t] = -xi[t];
              yr[t] = S.g.br[kr][32 * wc + 16 * t + (lane & 15)];
              yi[t] = S.g.bi[kr][32 * wc + 16 * t + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) {
                cr[a][b] = Mma<double>::fma(xr[a], yr[b], cr[a][b]);
                ci[a][b] = Mma<double>::fma(xr[a], yi[b], ci[a][b]);
              }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
              for (int b = 0; b < 2; b++) {
                cr[a][b] = Mma<double>::fma(nxi[a], yi[b], cr[a][b]);
                ci[a][b] = Mma<double>::fma(xi[a], yr[b], ci[a][b]);
              }
          }
        }
      }
      if (active) {
        double* ob = out + 2 * (size_t)i0 * kSynthB;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
              const int i = 32 * wr + 16 * a + Mma<double>::row(lane, e);
              const int j = 32 * wc + 16 * b + (lane & 15);
              if (i < rows) {
                const size_t t = (size_t)i * kSynthB + j;
                ob[2 * t] = cr[a][b][e];
                ob[2 * t + 1] = ci[a][b][e];
              }
            }
      }
    }
    // the points after the last full row: one term per thread, summed over the workgroup
    for (int64_t t = nfull * kSynthB; t < npts; t++) {
      double fr = 0.0, fi = 0.0;
      for (int k = tid; k < n; k += kThreads) {
        double sn, cs;
        sincos(lam[k] * (double)t, &sn, &cs);
        fr += gre[k] * cs + gim[k] * sn;
        fi += gim[k] * cs - gre[k] * sn;
      }
      fr = wave_sum_dpp(fr);
      fi = wave_sum_dpp(fi);
      if (lane == 0) {
        S.red[wave][0] = fr;
        S.red[wave][1] = fi;
      }
      __syncthreads();
      if (tid == 0) {
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < kThreads / 64; w++) {
          sr += S.red[w][0];
          si += S.red[w][1];
        }
        out[2 * t] = sr;
        out[2 * t + 1] = si;
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

}  // namespace spk

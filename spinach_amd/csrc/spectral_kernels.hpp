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
//              applied on the fly to conj(coil) and rho0. The 192 x 192 FP64
//              matrix lives in the registers of a 512-thread workgroup (each
//              thread 3 rows x 24 columns, cyclic), so the symmetric
//              matrix-vector products and rank-2 updates of every step run
//              from registers with the Householder vectors broadcast from LDS.
//   k_tql      implicit symmetric QR with Wilkinson shifts on T (one lane per
//              orientation), the Givens rotations applied to the two
//              transformed vectors only: the eigenvectors are never formed.
//   k_synth    f(64*i + j) = sum_k [g_k e^{-i lambda_k 64 i}] [e^{-i lambda_k j}]
//              as one complex matrix product per 64 x 64 block of time
//              points, on the matrix cores.
#pragma once

#include "powder_kernels.hpp"

namespace spk {

constexpr int kTriNP = 192;      // largest padded dimension of the register layout
constexpr int kTriThreads = 512;
constexpr int kTriRowsPer = 3;   // kTriNP / 64 row groups
constexpr int kTriColsPer = 24;  // kTriNP / 8 column groups
constexpr int kSpecStride = 6 * kTriNP;  // per item: d | e | four vectors
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
  const double* coil;
  const double* alphas;
  const double* betas;
  const double* gammas;
  double dt;
  int64_t nsteps;
  double* work;          // per-workgroup workspace (k_tridiag: dense np x np, all zero between uses;
                         // k_synth: U and W blocks)
  int64_t work_stride;   // doubles per workgroup
  double* spec;          // kSpecStride doubles per spectral item
  int32_t* status;       // per item: QR sweeps used, or -1 if the iteration did not converge
  double* per_or;        // per-orientation FIDs, interleaved complex, (nsteps+1) per orientation
};

// Scatter (sign = +1) or clear (sign = 0) the real symmetric generator
// A = dt*(x + x^T)/2 of one orientation into a dense row-major np x np buffer.
__device__ inline void scatter_real_generator(const SpecParams& p, const SysDesc& sd, const double* cre,
                                              const double* cim, double* W, bool clear) {
  const int32_t* prow = p.prow + sd.pat_off;
  const int32_t* pcol = p.pcol + sd.pat_off;
  const int32_t* ptp = p.ptpos + sd.pat_off;
  const double* vals = p.vals + 2 * sd.val_off;
  const int32_t* slots = p.slots + sd.slot_off;
  const int nnz = sd.nnz, np = sd.np;
  for (int e = threadIdx.x; e < nnz; e += blockDim.x) {
    const size_t off = (size_t)prow[e] * np + pcol[e];
    if (clear) {
      W[off] = 0.0;
      continue;
    }
    const int et = ptp[e];
    double xr = vals[2 * e], tr = vals[2 * et];
    for (int q = 0; q < sd.nq; q++) {
      const double* qv = vals + 2 * (size_t)(q + 1) * nnz;
      const double cr = cre[slots[q]], ci = cim[slots[q]];
      xr += cr * qv[2 * e] - ci * qv[2 * e + 1];
      tr += cr * qv[2 * et] - ci * qv[2 * et + 1];
    }
    W[off] = p.dt * 0.5 * (xr + tr);
  }
}

// ------------------------------------------------------------ k_tridiag
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

// Row jj of the register-resident matrix into xb and the sum of squares of
// its entries right of the first subdiagonal position into *sig. The owners
// are the eight threads of row group jj % 64 (one aligned 8-lane group);
// every thread calls this (the DPP sum needs the whole wave active). Indices
// into R are compile-time constants, so R stays in registers.
__device__ __forceinline__ void extract_row(const double (&R)[kTriRowsPer][kTriColsPer], int jj, int tr, int tc,
                                            double* xb, double* sig) {
  double part = 0.0;
#pragma unroll
  for (int a = 0; a < kTriRowsPer; a++) {
    if (tr + 64 * a == jj) {
#pragma unroll
      for (int b = 0; b < kTriColsPer; b++) {
        const double v = R[a][b];
        xb[tc + 8 * b] = v;
        if (tc + 8 * b >= jj + 2) part = fma(v, v, part);
      }
    }
  }
  part = sum8(part);
  if (tr == (jj & 63) && tc == 0) *sig = part;
}

__global__ void __launch_bounds__(kTriThreads, 1) k_tridiag(SpecParams p) {
  __shared__ double sx[2][kTriNP];  // rows j (Householder source), double-buffered
  __shared__ double ssig[2];
  __shared__ double sp[kTriNP];     // p = tau*A*v
  __shared__ double sq[kTriThreads / 64];  // per-wave parts of v'p
  __shared__ double su[4][kTriNP];
  __shared__ double sd[kTriNP], se[kTriNP];
  __shared__ double cre[34], cim[34];
  __shared__ int sitem;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int tr = tid >> 3, tc = tid & 7;  // row group 0..63, column group 0..7
  double* W = p.work + (size_t)blockIdx.x * p.work_stride;

  for (;;) {
    if (tid == 0) sitem = (int)atomicAdd(&p.queue[0], 1u);
    __syncthreads();
    const int64_t it = sitem;
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sdesc = p.sys[item.sys];
    const int n = sdesc.n, np = sdesc.np;
    const int64_t o = item.orient;
    if (tid == 0) wigner_coeffs(p.alphas[o], p.betas[o], p.gammas[o], cre, cim);
    __syncthreads();
    scatter_real_generator(p, sdesc, cre, cim, W, false);
    {
      const double* coilv = p.coil + 2 * sdesc.vec_off;
      const double* rhov = p.rho0 + 2 * sdesc.vec_off;
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
    double R[kTriRowsPer][kTriColsPer];
#pragma unroll
    for (int a = 0; a < kTriRowsPer; a++)
#pragma unroll
      for (int b = 0; b < kTriColsPer; b++) {
        const int i = tr + 64 * a, j = tc + 8 * b;
        R[a][b] = (i < np && j < np) ? W[(size_t)i * np + j] : 0.0;
      }
    __syncthreads();
    // restore the all-zero workspace for the next item
    scatter_real_generator(p, sdesc, cre, cim, W, true);
    extract_row(R, 0, tr, tc, sx[0], &ssig[0]);
    __syncthreads();

    for (int j = 0; j < n - 2; j++) {
      const int cb = j & 1;
      const double* x = sx[cb];
      // Householder vector v (v_{j+1} = x_1 - alpha, v_i = x_i below), tau = 2/v'v,
      // formed redundantly by every thread from row j
      const double sigma = ssig[cb];
      const double x1 = x[j + 1];
      double tau = 0.0, alpha = x1, v1 = 0.0;
      if (sigma > 0.0) {
        const double nrm = sqrt(x1 * x1 + sigma);
        alpha = (x1 > 0.0) ? -nrm : nrm;
        v1 = x1 - alpha;
        tau = 2.0 / (v1 * v1 + sigma);
      }
      if (tid == 0) {
        sd[j] = x[j];
        se[j] = alpha;
      }
      if (tau != 0.0) {  // uniform
        auto vat = [&](int i) -> double { return (i == j + 1) ? v1 : ((i > j + 1) ? x[i] : 0.0); };
        // p = tau*A*v on the trailing rows and the per-wave parts of v'p
        double part[kTriRowsPer] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int b = 0; b < kTriColsPer; b++) {
          if (8 * b + 7 > j) {
            const double vb = vat(tc + 8 * b);
#pragma unroll
            for (int a = 0; a < kTriRowsPer; a++)
              if (64 * a + 8 * wave + 7 > j) part[a] = fma(R[a][b], vb, part[a]);
          }
        }
        double qsum = 0.0;
#pragma unroll
        for (int a = 0; a < kTriRowsPer; a++) {
          const int i = tr + 64 * a;
          double s = sum8(part[a]) * tau;
          if (i <= j) s = 0.0;
          if (tc == 0) sp[i] = s;
          qsum = fma(vat(i), s, qsum);
        }
        // the eight lanes of a group hold the same qsum
        const double qw = wave_sum_dpp(qsum) * 0.125;
        if (lane == 0) sq[wave] = qw;
        // Q^T applied to conj(coil) and rho0, one vector per wave: u <- u - tau*v*(v'u)
        if (wave < 4) {
          double s = 0.0;
          for (int i = j + 1 + lane; i < n; i += 64) s = fma(vat(i), su[wave][i], s);
          s = wave_sum_dpp(s) * tau;
          for (int i = j + 1 + lane; i < n; i += 64) su[wave][i] -= s * vat(i);
        }
        __syncthreads();
        // A <- A - v w' - w v' with w = p - (tau/2)(v'p) v
        double qt = 0.0;
#pragma unroll
        for (int w8 = 0; w8 < kTriThreads / 64; w8++) qt += sq[w8];
        const double K = 0.5 * tau * qt;
        double vi[kTriRowsPer], wi[kTriRowsPer];
#pragma unroll
        for (int a = 0; a < kTriRowsPer; a++) {
          const int i = tr + 64 * a;
          vi[a] = vat(i);
          wi[a] = sp[i] - K * vi[a];
        }
#pragma unroll
        for (int b = 0; b < kTriColsPer; b++) {
          if (8 * b + 7 > j) {
            const double vb = vat(tc + 8 * b), wb = sp[tc + 8 * b] - K * vb;
#pragma unroll
            for (int a = 0; a < kTriRowsPer; a++)
              if (64 * a + 8 * wave + 7 > j) R[a][b] = fma(-vi[a], wb, fma(-wi[a], vb, R[a][b]));
          }
        }
      }
      // the owners of row j+1 publish it (after their own update) for the next step
      extract_row(R, j + 1, tr, tc, sx[cb ^ 1], &ssig[cb ^ 1]);
      __syncthreads();
    }
    // trailing 2 x 2 (or 1 x 1) block: row max(n-2, 0) is in its buffer; publish row n-1
    if (n >= 2) {
      extract_row(R, n - 1, tr, tc, sx[(n - 1) & 1], &ssig[(n - 1) & 1]);
      __syncthreads();
      if (tid == 0) {
        sd[n - 2] = sx[(n - 2) & 1][n - 2];
        se[n - 2] = sx[(n - 2) & 1][n - 1];
        sd[n - 1] = sx[(n - 1) & 1][n - 1];
      }
    } else if (n == 1) {
      if (tid == 0) sd[0] = sx[0][0];
    }
    __syncthreads();
    double* out = p.spec + (size_t)it * kSpecStride;
    for (int i = tid; i < kTriNP; i += kTriThreads) {
      out[i] = sd[i];
      out[kTriNP + i] = se[i];
#pragma unroll
      for (int q = 0; q < 4; q++) out[(2 + q) * kTriNP + i] = su[q][i];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ k_tql
// Implicit symmetric tridiagonal QR with Wilkinson shifts (Golub & Van Loan,
// Matrix Computations, algorithms 8.3.2/8.3.3), deflating at |e_i| <= eps*|T|
// (a normwise backward-stable test). Each Givens rotation is applied to the
// four real vectors (conj(coil) and rho0, real and imaginary parts) instead
// of an eigenvector matrix. One lane per orientation: the rotations form a
// dependent chain, so its operands live in LDS (kTqlLanes orientations per
// workgroup, rows padded so that lanes spread over the banks). A chain is
// bound by instruction issue (~100 FP64 instructions per rotation), so the
// orientations of a workgroup share one wave (measured on MI355X for the
// 185-state case: 16 per wave 14.6 ms, 4 waves x 4 25.5 ms, 16 waves x 1
// 27.2 ms for 4000 orientations). On exit the item's d holds the eigenvalues
// and its first two vectors hold g_k = (coil' v_k)(v_k^T rho0).
constexpr int kTqlPerWave = 16;
constexpr int kTqlWaves = 1;
constexpr int kTqlLanes = kTqlPerWave * kTqlWaves;  // orientations per workgroup
constexpr int kTqlStride = 6 * kTriNP + 2;          // doubles per orientation

__global__ void __launch_bounds__(64 * kTqlWaves) k_tql(SpecParams p) {
  __shared__ double lds[kTqlLanes * kTqlStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane >= kTqlPerWave) return;
  const int slot = wave * kTqlPerWave + lane;
  const int64_t it = (int64_t)blockIdx.x * kTqlLanes + slot;
  if (it >= p.n_items) return;
  const int n = p.sys[p.items[it].sys].n;
  double* gsrc = p.spec + (size_t)it * kSpecStride;
  double* d = lds + slot * kTqlStride;
  double* e = d + kTriNP;
  double* u0 = d + 2 * kTriNP;
  double* u1 = d + 3 * kTriNP;
  double* u2 = d + 4 * kTriNP;
  double* u3 = d + 5 * kTriNP;
  for (int q = 0; q < 6; q++)
    for (int i = 0; i < n; i++) d[q * kTriNP + i] = gsrc[q * kTriNP + i];
  double anorm = 0.0;
  for (int i = 0; i < n; i++) {
    double r = fabs(d[i]);
    if (i + 1 < n) r += fabs(e[i]);
    if (i > 0) r += fabs(e[i - 1]);
    anorm = fmax(anorm, r);
  }
  const double tol = 2.220446049250313e-16 * anorm;
  int hi = n - 1;
  int sweeps = 0;
  bool failed = false;
  while (hi > 0) {
    if (fabs(e[hi - 1]) <= tol) {
      e[hi - 1] = 0.0;
      hi--;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && fabs(e[lo - 1]) > tol) lo--;
    if (++sweeps > 40 * n) {
      failed = true;
      break;
    }
    // Wilkinson shift from the trailing 2 x 2 block
    const double dd = 0.5 * (d[hi - 1] - d[hi]), bb = e[hi - 1];
    const double h = sqrt(dd * dd + bb * bb);
    const double mu = d[hi] - bb * bb / (dd + (dd >= 0.0 ? h : -h));
    double x = d[lo] - mu, z = e[lo];
    double dk = d[lo], ek = e[lo];
    double dk1 = d[lo + 1], ek1 = (lo + 1 < hi) ? e[lo + 1] : 0.0;
    double a0 = u0[lo], a1 = u1[lo], a2 = u2[lo], a3 = u3[lo];
    double b0 = u0[lo + 1], b1 = u1[lo + 1], b2 = u2[lo + 1], b3 = u3[lo + 1];
    for (int k = lo; k < hi; k++) {
      // operands of the next rotation, loaded ahead of the dependent chain
      const bool more = (k + 1 < hi);
      double nd = 0.0, ne = 0.0, n0 = 0.0, n1 = 0.0, n2 = 0.0, n3 = 0.0;
      if (more) {
        nd = d[k + 2];
        ne = (k + 2 < hi) ? e[k + 2] : 0.0;
        n0 = u0[k + 2];
        n1 = u1[k + 2];
        n2 = u2[k + 2];
        n3 = u3[k + 2];
      }
      const double r2 = x * x + z * z;
      double c = 1.0, s = 0.0, r = 0.0;
      if (r2 > 0.0) {
        const double ri = rsqrt(r2);
        c = x * ri;
        s = z * ri;
        r = r2 * ri;
      }
      if (k > lo) e[k - 1] = r;
      const double cc = c * c, ss = s * s, cs = c * s;
      const double dkn = cc * dk + 2.0 * cs * ek + ss * dk1;
      const double dk1n = ss * dk - 2.0 * cs * ek + cc * dk1;
      const double ekn = cs * (dk1 - dk) + (cc - ss) * ek;
      d[k] = dkn;
      u0[k] = c * a0 + s * b0;
      u1[k] = c * a1 + s * b1;
      u2[k] = c * a2 + s * b2;
      u3[k] = c * a3 + s * b3;
      a0 = c * b0 - s * a0;
      a1 = c * b1 - s * a1;
      a2 = c * b2 - s * a2;
      a3 = c * b3 - s * a3;
      if (more) {
        x = ekn;
        z = s * ek1;
        ek = c * ek1;
        dk = dk1n;
        dk1 = nd;
        ek1 = ne;
        b0 = n0;
        b1 = n1;
        b2 = n2;
        b3 = n3;
      } else {
        d[k + 1] = dk1n;
        e[k] = ekn;
        u0[k + 1] = a0;
        u1[k + 1] = a1;
        u2[k + 1] = a2;
        u3[k + 1] = a3;
      }
    }
  }
  // eigenvalues and g_k = a_k * b_k (a = transformed conj(coil), b = transformed rho0)
  for (int k = 0; k < n; k++) {
    const double ar = u0[k], ai = u1[k], br = u2[k], bi = u3[k];
    gsrc[k] = d[k];
    gsrc[2 * kTriNP + k] = ar * br - ai * bi;
    gsrc[3 * kTriNP + k] = ar * bi + ai * br;
  }
  p.status[it] = failed ? -1 : sweeps;
}

// ------------------------------------------------------------ k_synth
__global__ void __launch_bounds__(kThreads, 2) k_synth(SpecParams p) {
  __shared__ GemmLds<double> g;
  __shared__ int sitem;
  const int tid = threadIdx.x;
  double* ws = p.work + (size_t)blockIdx.x * p.work_stride;
  const int64_t npts = p.nsteps + 1;
  const int64_t nfull = npts / kSynthB;
  for (;;) {
    if (tid == 0) sitem = (int)atomicAdd(&p.queue[1], 1u);
    __syncthreads();
    const int64_t it = sitem;
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sdesc = p.sys[item.sys];
    const int n = sdesc.n, np = sdesc.np;
    const double* lam = p.spec + (size_t)it * kSpecStride;
    const double* gre = lam + 2 * kTriNP;
    const double* gim = lam + 3 * kTriNP;
    double* out = p.per_or + 2 * (size_t)item.orient * npts;
    // W[k][j] = exp(-i lambda_k j), j < 64 (rows k >= n are zero)
    double* Wre = ws;
    double* Wim = ws + (size_t)np * kSynthB;
    double* Ure = ws + 2 * (size_t)np * kSynthB;
    double* Uim = Ure + (size_t)kSynthB * np;
    for (int e = tid; e < np * kSynthB; e += kThreads) {
      const int k = e / kSynthB, j = e % kSynthB;
      double c = 0.0, s = 0.0;
      if (k < n) sincos(lam[k] * (double)j, &s, &c);
      Wre[e] = c;
      Wim[e] = -s;
    }
    for (int64_t i0 = 0; i0 < nfull; i0 += kSynthB) {
      const int rows = (int)((nfull - i0 < kSynthB) ? (nfull - i0) : kSynthB);
      __syncthreads();
      // U[i][k] = g_k exp(-i lambda_k 64 (i0 + i))
      for (int e = tid; e < rows * np; e += kThreads) {
        const int i = e / np, k = e % np;
        double ur = 0.0, ui = 0.0;
        if (k < n) {
          double s, c;
          sincos(lam[k] * (double)(kSynthB * (i0 + i)), &s, &c);
          ur = gre[k] * c + gim[k] * s;
          ui = gim[k] * c - gre[k] * s;
        }
        Ure[e] = ur;
        Uim[e] = ui;
      }
      __syncthreads();
      double* ob = out + 2 * (size_t)i0 * kSynthB;
      wg_gemm_t<double, false, false, false, false>(rows, kSynthB, np, Ure, Uim, np, Wre, Wim, kSynthB,
                                                    [=](int i, int j, double vr, double vi) {
                                                      const size_t t = (size_t)i * kSynthB + j;
                                                      ob[2 * t] = vr;
                                                      ob[2 * t + 1] = vi;
                                                    },
                                                    g);
    }
    // the points after the last full block
    for (int64_t t = nfull * kSynthB + tid; t < npts; t += kThreads) {
      double fr = 0.0, fi = 0.0;
      for (int k = 0; k < n; k++) {
        double s, c;
        sincos(lam[k] * (double)t, &s, &c);
        fr += gre[k] * c + gim[k] * s;
        fi += gim[k] * c - gre[k] * s;
      }
      out[2 * t] = fr;
      out[2 * t + 1] = fi;
    }
    __syncthreads();
  }
}

}  // namespace spk

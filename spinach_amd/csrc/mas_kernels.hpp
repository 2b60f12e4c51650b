// Magic-angle-spinning powder propagation in Hilbert space (dense N x N,
// N = 2^nspins), one workgroup per orientation.
//
// Rotor composition (kernel/utilities/rotor_stack.m:93-147): at rotor phase
// phi the rank-r Wigner matrix is
//     D(phi) = D_rot2lab * diag(exp(-i m phi)) * D_lab2rot * D_initial(Omega)
// ('magnet' frame; the 'rotor' frame drops D_lab2rot), with phases
// phi_j = 2*pi*j/ticks (kernel/derivatives/fourdif.m:38), and
//     H_j = I + sum_{r,k,m} D(phi_j)(k,m) Q{r}{k,m}.
// Each tick is propagated by P_j = expm(-i H_j dt) (Hilbert branch of
// kernel/propagator.m:40-42), rho -> P rho P' (kernel/step.m:115-137), and
// detected with hdot(coil, rho) = sum(conj(coil).*rho) (utilities/hdot.m:23-29).
//
// Evaluation order on the GPU (same mathematics): with U_j = P_{j-1}...P_0
// and the period propagator U = U_ticks,
//     f(k*ticks + j) = hdot(C_j, rho_k),  C_j = U_j' * coil * U_j,
//     rho_k = U^k rho0 U'^k,
// so one rotor period of expm and products gives every C_j, the rho chain
// needs two products per period, and all detections are one matrix product
// conj(C) * rho^T over the N^2 flattened elements.
#pragma once

#include "powder_kernels.hpp"

namespace spk {

template <typename R>
struct MasParams {
  const int32_t* prow;
  const int32_t* pcol;
  int32_t nnz;
  int32_t nq;
  const double* vals;    // (1+nq)*nnz interleaved complex: I, then Q slots
  const int32_t* slots;  // Q slot ids
  const double* rho0;    // np x np row-major interleaved complex
  const double* coil;    // np x np row-major interleaved complex
  const double* alphas;
  const double* betas;
  const double* gammas;
  int64_t n_or;
  const double* rot2lab;  // 9 + 25 interleaved complex (rank 1, rank 2)
  const double* lab2rot;  // 9 + 25 interleaved complex (identity in the rotor frame)
  double dt;
  int32_t ticks;
  int32_t n_rho;
  int64_t nsteps;
  int32_t np;
  int32_t taylor_k;
  int32_t ps_q;
  double theta;
  R* work;
  int64_t work_stride;
  unsigned int* queue;
  double* per_or;
  int32_t* sq_out;  // per orientation: largest scaling exponent over the ticks
};

template <typename R>
struct MasShared {
  GemmLds<R> g;
  double B1r[9], B1i[9], B2r[25], B2i[25];  // D_lab2rot * D_initial
  double cre[34], cim[34];
  double red[4];
  double tc[32];
  int item;
  int sexp;
  int smax;
};

// c[k][m] = sum_j A(k,j) exp(-i m_j phi) B(j,m) for one rank (size sz)
__device__ inline void rotor_coeffs(int sz, const double* A, const double* Br, const double* Bi, double phi,
                                    double* cre, double* cim) {
  int l = (sz - 1) / 2;
  for (int k = 0; k < sz; k++)
    for (int m = 0; m < sz; m++) {
      double sr = 0, si = 0;
      for (int j = 0; j < sz; j++) {
        int mj = l - j;
        double er = cos(mj * phi), ei = -sin(mj * phi);
        double ar = A[2 * (k * sz + j)], ai = A[2 * (k * sz + j) + 1];
        // a * e
        double tr = ar * er - ai * ei, ti = ar * ei + ai * er;
        double br = Br[j * sz + m], bi = Bi[j * sz + m];
        sr += tr * br - ti * bi;
        si += tr * bi + ti * br;
      }
      cre[k * sz + m] = sr;
      cim[k * sz + m] = si;
    }
}

template <typename R>
__global__ void __launch_bounds__(kThreads) k_mas(MasParams<R> p) {
  __shared__ MasShared<R> sh;
  const int tid = threadIdx.x;
  const int np = p.np;
  const size_t pl = (size_t)np * np;
  const size_t mat = 2 * pl;
  R* ws = p.work + (size_t)blockIdx.x * p.work_stride;
  Buf<R> bf[6];
  for (int b = 0; b < 6; b++) bf[b] = Buf<R>{ws + b * mat, pl, np};
  Buf<R> U{ws + 6 * mat, pl, np}, Ut{ws + 7 * mat, pl, np}, T{ws + 8 * mat, pl, np}, Cl{ws + 9 * mat, pl, np};
  R* Cb = ws + 10 * mat;                       // ticks x np^2, planar
  R* Rb = Cb + 2 * (size_t)p.ticks * pl;       // n_rho x np^2, planar
  const size_t cpl = (size_t)p.ticks * pl;     // plane size of Cb
  const size_t rpl = (size_t)p.n_rho * pl;     // plane size of Rb

  if (tid < 32) {
    double f = 1.0;
    for (int k = 2; k <= tid; k++) f *= k;
    sh.tc[tid] = 1.0 / f;
  }
  // coil in working precision, once per workgroup
  for (size_t e = tid; e < pl; e += kThreads) {
    Cl.re()[e] = (R)p.coil[2 * e];
    Cl.im()[e] = (R)p.coil[2 * e + 1];
  }
  __syncthreads();

  for (;;) {
    if (tid == 0) sh.item = (int)atomicAdd(p.queue, 1u);
    __syncthreads();
    const int64_t o = sh.item;
    __syncthreads();
    if (o >= p.n_or) break;
    if (tid == 0) {
      // B = D_lab2rot * D_initial for ranks 1 and 2
      double dr[34], di[34];
      wigner_coeffs(p.alphas[o], p.betas[o], p.gammas[o], dr, di);
      for (int k = 0; k < 3; k++)
        for (int m = 0; m < 3; m++) {
          double sr = 0, si = 0;
          for (int j = 0; j < 3; j++) {
            double ar = p.lab2rot[2 * (k * 3 + j)], ai = p.lab2rot[2 * (k * 3 + j) + 1];
            double br = dr[j * 3 + m], bi = di[j * 3 + m];
            sr += ar * br - ai * bi;
            si += ar * bi + ai * br;
          }
          sh.B1r[k * 3 + m] = sr;
          sh.B1i[k * 3 + m] = si;
        }
      for (int k = 0; k < 5; k++)
        for (int m = 0; m < 5; m++) {
          double sr = 0, si = 0;
          for (int j = 0; j < 5; j++) {
            double ar = p.lab2rot[2 * (9 + k * 5 + j)], ai = p.lab2rot[2 * (9 + k * 5 + j) + 1];
            double br = dr[9 + j * 5 + m], bi = di[9 + j * 5 + m];
            sr += ar * br - ai * bi;
            si += ar * bi + ai * br;
          }
          sh.B2r[k * 5 + m] = sr;
          sh.B2i[k * 5 + m] = si;
        }
      sh.smax = 0;
    }
    // U = I, C_0 = coil
    for (size_t e = tid; e < pl; e += kThreads) {
      U.re()[e] = (e / np == e % np) ? R(1) : R(0);
      U.im()[e] = R(0);
      Cb[e] = Cl.re()[e];
      Cb[cpl + e] = Cl.im()[e];
    }
    __syncthreads();
    for (int j = 0; j < p.ticks; j++) {
      if (j > 0) {
        // C_j = U' * (coil * U)
        R* tre = T.re();
        R* tim = T.im();
        wg_gemm_nn<R>(np, np, np, Cl.re(), Cl.im(), np, U.re(), U.im(), np,
                   [=](int a, int b, R vr, R vi) {
                     tre[(size_t)a * np + b] = vr;
                     tim[(size_t)a * np + b] = vi;
                   },
                   sh.g);
        R* cre = Cb + (size_t)j * pl;
        R* cim = Cb + cpl + (size_t)j * pl;
        wg_gemm_t<R, true, true, false, false>(np, np, np, U.re(), U.im(), np, T.re(), T.im(), np,
                   [=](int a, int b, R vr, R vi) {
                     cre[(size_t)a * np + b] = vr;
                     cim[(size_t)a * np + b] = vi;
                   },
                   sh.g);
      }
      // generator at rotor phase phi_j
      if (tid == 0) {
        const double phi = 6.283185307179586476925 * (double)j / (double)p.ticks;
        rotor_coeffs(3, p.rot2lab, sh.B1r, sh.B1i, phi, sh.cre, sh.cim);
        rotor_coeffs(5, p.rot2lab + 18, sh.B2r, sh.B2i, phi, sh.cre + 9, sh.cim + 9);
      }
      R* X = bf[0].p;
      for (size_t e = tid; e < mat; e += kThreads) X[e] = R(0);
      __syncthreads();
      for (int e = tid; e < p.nnz; e += kThreads) {
        double xr = p.vals[2 * e], xi = p.vals[2 * e + 1];
        for (int qq = 0; qq < p.nq; qq++) {
          const double* qv = p.vals + 2 * (size_t)(qq + 1) * p.nnz;
          double cr = sh.cre[p.slots[qq]], ci = sh.cim[p.slots[qq]];
          xr += cr * qv[2 * e] - ci * qv[2 * e + 1];
          xi += cr * qv[2 * e + 1] + ci * qv[2 * e];
        }
        size_t off = (size_t)p.prow[e] * np + p.pcol[e];
        X[off] = (R)(p.dt * xi);        // -i*dt*H
        X[pl + off] = (R)(-p.dt * xr);
      }
      __syncthreads();
      {
        double colmax = 0.0;
        for (int c = tid; c < np; c += kThreads) {
          double sacc = 0.0;
          for (int i = 0; i < np; i++) {
            double a = (double)X[(size_t)i * np + c], b = (double)X[pl + (size_t)i * np + c];
            sacc += sqrt(a * a + b * b);
          }
          colmax = fmax(colmax, sacc);
        }
        double nrm = wg_max(colmax, sh.red);
        if (tid == 0) {
          int sx = 0;
          if (nrm > p.theta) sx = (int)ceil(log2(nrm / p.theta));
          sh.sexp = sx;
          sh.smax = sx > sh.smax ? sx : sh.smax;
        }
        __syncthreads();
      }
      const int sexp = sh.sexp;
      {
        const R scl = (R)ldexp(1.0, -sexp);
        for (size_t e = tid; e < mat; e += kThreads) X[e] *= scl;
        __syncthreads();
      }
      Buf<R> P = wg_expm<R>(bf, sexp, p.ps_q, p.taylor_k, sh.tc, sh.g);
      // U <- P * U
      {
        R* ure = Ut.re();
        R* uim = Ut.im();
        wg_gemm_nn<R>(np, np, np, P.re(), P.im(), np, U.re(), U.im(), np,
                   [=](int a, int b, R vr, R vi) {
                     ure[(size_t)a * np + b] = vr;
                     uim[(size_t)a * np + b] = vi;
                   },
                   sh.g);
        Buf<R> t2 = U;
        U = Ut;
        Ut = t2;
      }
    }
    if (tid == 0 && p.sq_out) p.sq_out[o] = sh.smax;
    // rho chain: rho_0 = rho0, rho_k = U rho_{k-1} U'
    for (size_t e = tid; e < pl; e += kThreads) {
      Rb[e] = (R)p.rho0[2 * e];
      Rb[rpl + e] = (R)p.rho0[2 * e + 1];
    }
    __syncthreads();
    for (int k = 1; k < p.n_rho; k++) {
      const R* prev_re = Rb + (size_t)(k - 1) * pl;
      const R* prev_im = Rb + rpl + (size_t)(k - 1) * pl;
      R* tre = T.re();
      R* tim = T.im();
      wg_gemm_nn<R>(np, np, np, U.re(), U.im(), np, prev_re, prev_im, np,
                 [=](int a, int b, R vr, R vi) {
                   tre[(size_t)a * np + b] = vr;
                   tim[(size_t)a * np + b] = vi;
                 },
                 sh.g);
      R* rre = Rb + (size_t)k * pl;
      R* rim = Rb + rpl + (size_t)k * pl;
      wg_gemm_t<R, false, false, true, true>(np, np, np, T.re(), T.im(), np, U.re(), U.im(), np,
                 [=](int a, int b, R vr, R vi) {
                   rre[(size_t)a * np + b] = vr;
                   rim[(size_t)a * np + b] = vi;
                 },
                 sh.g);
    }
    // f(k*ticks + j) = sum conj(C_j) .* rho_k
    {
      const int64_t npts = p.nsteps + 1;
      double* out = p.per_or + 2 * (size_t)o * npts;
      const int ticks = p.ticks;
      wg_gemm_t<R, false, true, true, false>(ticks, p.n_rho, (int)pl, Cb, Cb + cpl, (int)pl, Rb, Rb + rpl,
                                             (int)pl,
                 [=](int jj, int kk, R vr, R vi) {
                   int64_t t = (int64_t)kk * ticks + jj;
                   if (t < npts) {
                     out[2 * t] = (double)vr;
                     out[2 * t + 1] = (double)vi;
                   }
                 },
                 sh.g);
    }
  }
}

}  // namespace spk

namespace spk {

// ------------------------------------------------------------------ tiled MAS (FP64)
// The secular Hamiltonian of a system with several species conserves each species' total Fz,
// so H(phi) = I + sum D(phi)(k,m) Q{r}{k,m} is block diagonal at every rotor phase: the
// connected components of the union pattern of I and the Q slots are invariant subspaces of
// every tick propagator and of every U_j. When no component exceeds 16 states, the planner packs
// the components into tiles of 16 (first-fit decreasing) and numbers the basis tile by tile;
// then H(phi), each P_j and each U_j are block diagonal in 16 x 16 tiles, every product of the
// dense algorithm above splits into independent 16 x 16 products, and
//     C_j = U_j' coil U_j,  rho_k = U^k rho0 U'^k
// need only the tile pairs (s, t) where both coil and rho0 have entries (the detection
// sum(conj(C_j) .* rho_k) sees nothing else); inside a pair, only the elements of the component
// blocks where both have entries are kept (emap), so the detection product runs over those.
// Same mathematics, same evaluation order of the Taylor polynomial; one wave owns a tile (tiles
// wave, wave + 4, ...) and a 16 x 16 complex product is 16 FP64 MFMAs on registers. The
// generator is assembled from its Fourier components in the rotor phase (formed once per
// orientation), so a tick reads five dense tiles instead of the sparse pattern and slot values.
//
// Register layouts of a 16 x 16 matrix in a wave (the MFMA f64 16x16x4 operand layouts):
//   C layout (accumulator):  lane l holds M[4e + l/16][l%16], e = 0..3;
//   A operand, k-chunk s:    lane l holds M[l%16][4s + l/16]  (the C layout of M');
//   B operand, k-chunk s:    lane l holds M[4s + l/16][l%16]  (the C layout itself, e = s).
// A matrix in the C layout therefore serves as a B operand as it is; as an A operand it is
// transposed through the wave's LDS scratch tile, or read transposed from LDS where it is stored.

constexpr int kTileN = 16;
constexpr int kTileLd = 17;                 // LDS row stride (one pad double: transposed reads spread over banks)
constexpr int kTilePlane = kTileN * kTileLd;  // doubles per real plane of an LDS tile
constexpr int kMasMaxTiles = 8;

struct MasTParams {
  const int32_t* epos;  // per pattern entry (sorted by tile): tile-local position r * 16 + c
  const int32_t* toff;  // per tile: first entry; ntile + 1 values
  int32_t ntile;
  int32_t nnz;
  int32_t nq;
  const double* vals;    // (1+nq)*nnz interleaved complex, entry order as epos: I, then Q slots
  const int32_t* slots;  // Q slot ids
  int32_t npair;
  int32_t coil_real;      // every coil tile is real: W = coil_st U_t takes two real products, not four
  const int32_t* pair_s;  // tile pairs where both coil and rho0 have entries
  const int32_t* pair_t;
  const double* coil_t;  // npair x 256 interleaved complex, the tile row-major
  const double* rho0_t;  // npair x 256 interleaved complex
  const int32_t* emap;   // npair x 256: compact index of a detected element, -1 elsewhere
  int32_t ldet;          // detected elements, padded to a multiple of 16
  const double* alphas;
  const double* betas;
  const double* gammas;
  int64_t n_or;
  const double* rot2lab;
  const double* lab2rot;
  double dt;
  int32_t ticks;
  int32_t n_rho;
  int64_t nsteps;
  int32_t taylor_k;
  int32_t ps_q;
  double theta;
  double* work;  // per workgroup: C (ticks x ldet) and rho (n_rho x ldet), planar complex, then the
                 // generator's Fourier components G_n (nnz x 5 complex, entry order as epos)
  int64_t work_stride;
  unsigned int* queue;
  double* per_or;
  int32_t* sq_out;
  int32_t* prod_out;  // per orientation: 16 x 16 real tile products made (a complex product counts 4)
};

struct MasTShared {
  union {
    GemmLds<double> g;  // the detection product, after the propagation
    struct {
      double U[kMasMaxTiles][2 * kTilePlane];     // the tiles of U_j, row-major planar
      double scr[kThreads / 64][2 * kTilePlane];  // per-wave scratch tile
    } t;
  } u;
  double tc[32];
  int item;
  int smax;
  int nprod;
};

struct CTile {  // a 16 x 16 complex matrix in the C layout
  Mma<double>::v4 r, i;
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// D = A * B: A as its A-operand chunks (ar, ai), B in the C layout
__device__ __forceinline__ CTile cmm16(const double (&ar)[4], const double (&ai)[4], const CTile& b) {
  typedef Mma<double>::v4 v4;
  v4 dr = v4{0, 0, 0, 0}, di = v4{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 4; s++) {
    dr = Mma<double>::fma(ar[s], b.r[s], dr);
    di = Mma<double>::fma(ar[s], b.i[s], di);
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    dr = Mma<double>::fma(-ai[s], b.i[s], dr);
    di = Mma<double>::fma(ai[s], b.r[s], di);
  }
  return CTile{dr, di};
}

// D = A * B with A real (its A-operand chunks ar), B complex in the C layout: 8 MFMAs
__device__ __forceinline__ CTile rmm16(const double (&ar)[4], const CTile& b) {
  typedef Mma<double>::v4 v4;
  v4 dr = v4{0, 0, 0, 0}, di = v4{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 4; s++) {
    dr = Mma<double>::fma(ar[s], b.r[s], dr);
    di = Mma<double>::fma(ar[s], b.i[s], di);
  }
  return CTile{dr, di};
}

// the A-operand chunks of M (in the C layout), through the wave's scratch tile
__device__ __forceinline__ void tile_aop(const CTile& m, double* scr, int lane, double (&ar)[4], double (&ai)[4]) {
  const int rq = lane >> 4, cq = lane & 15;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    scr[(4 * e + rq) * kTileLd + cq] = m.r[e];
    scr[kTilePlane + (4 * e + rq) * kTileLd + cq] = m.i[e];
  }
  wave_sync_lds();
#pragma unroll
  for (int s = 0; s < 4; s++) {
    ar[s] = scr[cq * kTileLd + 4 * s + rq];
    ai[s] = scr[kTilePlane + cq * kTileLd + 4 * s + rq];
  }
  wave_sync_lds();
}

// loads of a row-major planar LDS tile: the C layout, and the A-operand chunks (optionally of
// the conjugate)
__device__ __forceinline__ CTile tile_load_c(const double* t, int lane) {
  const int rq = lane >> 4, cq = lane & 15;
  CTile m;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    m.r[e] = t[(4 * e + rq) * kTileLd + cq];
    m.i[e] = t[kTilePlane + (4 * e + rq) * kTileLd + cq];
  }
  return m;
}

__device__ __forceinline__ void tile_load_aop(const double* t, int lane, bool conj, double (&ar)[4],
                                              double (&ai)[4]) {
  const int rq = lane >> 4, cq = lane & 15;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    ar[s] = t[cq * kTileLd + 4 * s + rq];
    const double x = t[kTilePlane + cq * kTileLd + 4 * s + rq];
    ai[s] = conj ? -x : x;
  }
}

__device__ __forceinline__ void tile_store(double* t, const CTile& m, int lane) {
  const int rq = lane >> 4, cq = lane & 15;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    t[(4 * e + rq) * kTileLd + cq] = m.r[e];
    t[kTilePlane + (4 * e + rq) * kTileLd + cq] = m.i[e];
  }
}

// a global tile (256 interleaved complex, row-major) in the C layout
__device__ __forceinline__ CTile tile_global_c(const double* g, int lane) {
  const int rq = lane >> 4, cq = lane & 15;
  CTile m;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    m.r[e] = g[2 * ((4 * e + rq) * kTileN + cq)];
    m.i[e] = g[2 * ((4 * e + rq) * kTileN + cq) + 1];
  }
  return m;
}

// acc += sum_{i2 < q, j0*q + i2 <= K} tc[j0*q + i2] X^i2 (X^0 = identity; x1, x2, x3 = X, X^2, X^3)
__device__ __forceinline__ void tile_taylor_block(CTile& acc, int j0, int q, int K, const double* tc, const CTile& x1,
                                                  const CTile& x2, const CTile& x3, int rq, int cq) {
#pragma unroll
  for (int i2 = 0; i2 < 4; i2++) {
    const int k = j0 * q + i2;
    if (i2 < q && k <= K) {
      const double c = tc[k];
      if (i2 == 0) {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (4 * e + rq == cq) acc.r[e] += c;
      } else {
        const CTile& x = (i2 == 1) ? x1 : ((i2 == 2) ? x2 : x3);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          acc.r[e] = fma(c, x.r[e], acc.r[e]);
          acc.i[e] = fma(c, x.i[e], acc.i[e]);
        }
      }
    }
  }
}

__global__ void __launch_bounds__(kThreads, 3) k_mas_tiled(MasTParams p) {
  typedef Mma<double>::v4 v4;
  __shared__ MasTShared sh;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rq = lane >> 4, cq = lane & 15;
  constexpr int kWaves = kThreads / 64;
  const int64_t L = p.ldet;
  double* Cre = p.work + (size_t)blockIdx.x * p.work_stride;  // ticks x L
  double* Cim = Cre + (size_t)p.ticks * L;
  double* Rre = Cim + (size_t)p.ticks * L;  // n_rho x L
  double* Rim = Rre + (size_t)p.n_rho * L;
  double* scr = sh.u.t.scr[wave];
  if (tid < 32) {
    double f = 1.0;
    for (int k = 2; k <= tid; k++) f *= k;
    sh.tc[tid] = 1.0 / f;
  }
  __syncthreads();
  const int K = p.taylor_k, q = p.ps_q;

  for (;;) {
    if (tid == 0) sh.item = (int)atomicAdd(p.queue, 1u);
    __syncthreads();
    const int64_t o = sh.item;
    __syncthreads();
    if (o >= p.n_or) break;
    // Setup of the orientation, one slot per thread, staged in the wave scratch tiles:
    // D = D_initial (34 Wigner coefficients), B = D_lab2rot * D per rank.
    // The slot coefficients are trigonometric polynomials of the rotor phase:
    //   c_(k,m)(phi) = sum_j A(k,j) exp(-i (l - j) phi) B(j,m) = sum_n a_(k,m),n exp(-i n phi),
    //   a_(k,m),n = A(k, l - n) B(l - n, m),  |n| <= l,
    // so X(phi) = -i dt H(phi) = sum_{n=-2..2} exp(-i n phi) G_n with G_n = -i dt (delta_n0 I +
    // sum_slots a_slot,n Q_slot), formed once per orientation.
    double* acr = &sh.u.t.scr[0][0];  // 34 slots x 5 (n + 2)
    double* aci = acr + 34 * 5;
    double* dre = &sh.u.t.scr[1][0];  // D by slot
    double* dim = dre + 34;
    double* bre = &sh.u.t.scr[2][0];  // B by slot
    double* bim = bre + 34;
    double* Gw = Rim + (size_t)p.n_rho * L;  // per pattern entry (tile order): G_n, n = -2..2
    if (tid < 34) wigner_coeff_slot(p.alphas[o], p.betas[o], p.gammas[o], tid, dre, dim);
    if (tid == 0) {
      sh.smax = 0;
      sh.nprod = 0;
    }
    __syncthreads();  // D written
    if (tid < 34) {
      const bool r1 = tid < 9;
      const int sz = r1 ? 3 : 5, base = r1 ? 0 : 9, km = tid - base, k = km / sz, m = km % sz;
      double sr = 0, si = 0;
      for (int j = 0; j < sz; j++) {
        const double ar = p.lab2rot[2 * (base + k * sz + j)], ai = p.lab2rot[2 * (base + k * sz + j) + 1];
        const double br = dre[base + j * sz + m], bi = dim[base + j * sz + m];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      bre[tid] = sr;
      bim[tid] = si;
    }
    __syncthreads();  // B written
    if (tid < 34 * 5) {
      const int slot = tid / 5, n = tid % 5 - 2;
      const bool r1 = slot < 9;
      const int l = r1 ? 1 : 2, sz = r1 ? 3 : 5, km = r1 ? slot : slot - 9;
      const int k = km / sz, m = km % sz, jn = l - n;
      double xr = 0.0, xi = 0.0;
      if (jn >= 0 && jn < sz) {
        const double* A = p.rot2lab + (r1 ? 0 : 18) + 2 * (k * sz + jn);
        const double br = bre[(r1 ? 0 : 9) + jn * sz + m], bi = bim[(r1 ? 0 : 9) + jn * sz + m];
        xr = A[0] * br - A[1] * bi;
        xi = A[0] * bi + A[1] * br;
      }
      acr[tid] = xr;
      aci[tid] = xi;
    }
    __syncthreads();  // a staged
    {
      const int ne = p.toff[p.ntile];
      for (int x = tid; x < ne * 5; x += kThreads) {
        const int e = x / 5, nn = x % 5;
        double hr = (nn == 2) ? p.vals[2 * e] : 0.0, hi = (nn == 2) ? p.vals[2 * e + 1] : 0.0;
        for (int qq = 0; qq < p.nq; qq++) {
          const double* qv = p.vals + 2 * (size_t)(qq + 1) * p.nnz;
          const double cr = acr[p.slots[qq] * 5 + nn], ci = aci[p.slots[qq] * 5 + nn];
          hr += cr * qv[2 * e] - ci * qv[2 * e + 1];
          hi += cr * qv[2 * e + 1] + ci * qv[2 * e];
        }
        Gw[10 * (size_t)e + 2 * nn] = p.dt * hi;  // -i dt H
        Gw[10 * (size_t)e + 2 * nn + 1] = -p.dt * hr;
      }
    }
    // U_0 = identity on every tile
    for (int t = wave; t < p.ntile; t += kWaves) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        sh.u.t.U[t][(4 * e + rq) * kTileLd + cq] = (4 * e + rq == cq) ? 1.0 : 0.0;
        sh.u.t.U[t][kTilePlane + (4 * e + rq) * kTileLd + cq] = 0.0;
      }
    }
    int smax = 0, nprod = 0;  // per wave
    for (int j = 0; j < p.ticks; j++) {
      __syncthreads();  // every tile of U_j stored (and, before the first tick, G written)
      // C_j = U_j' coil U_j on the tile pairs: W = coil_st U_t, C = U_s' W
      for (int pp = wave; pp < p.npair; pp += kWaves) {
        const int s = p.pair_s[pp], t = p.pair_t[pp];
        double car[4], cai[4];
        {
          const double* g = p.coil_t + (size_t)pp * 512;  // A operand of coil_st: element (cq, 4s' + rq)
#pragma unroll
          for (int s2 = 0; s2 < 4; s2++) {
            car[s2] = g[2 * (cq * kTileN + 4 * s2 + rq)];
            cai[s2] = g[2 * (cq * kTileN + 4 * s2 + rq) + 1];
          }
        }
        const CTile W = p.coil_real ? rmm16(car, tile_load_c(sh.u.t.U[t], lane))
                                    : cmm16(car, cai, tile_load_c(sh.u.t.U[t], lane));
        // A operand of U_s' (conjugate transpose): chunk s2, lane l: conj(U_s[4 s2 + l/16][l%16])
        double uar[4], uai[4];
        {
          const CTile us = tile_load_c(sh.u.t.U[s], lane);
#pragma unroll
          for (int s2 = 0; s2 < 4; s2++) {
            uar[s2] = us.r[s2];
            uai[s2] = -us.i[s2];
          }
        }
        const CTile C = cmm16(uar, uai, W);
        nprod += p.coil_real ? 6 : 8;
        double* cr = Cre + (size_t)j * L;
        double* ci = Cim + (size_t)j * L;
        const int32_t* em = p.emap + (size_t)pp * 256;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int x = em[(4 * e + rq) * kTileN + cq];
          if (x >= 0) {
            cr[x] = C.r[e];
            ci[x] = C.i[e];
          }
        }
      }
      __syncthreads();  // every wave has read the U tiles it needs for C_j
      const double phi = 6.283185307179586476925 * (double)j / (double)p.ticks;
      double s1, c1, s2, c2;
      sincos(phi, &s1, &c1);
      sincos(2.0 * phi, &s2, &c2);
      // exp(-i n phi), n = -2..2
      const double er[5] = {c2, c1, 1.0, c1, c2}, ei[5] = {s2, s1, 0.0, -s1, -s2};
      for (int t = wave; t < p.ntile; t += kWaves) {
        // X = sum_n exp(-i n phi) G_n on tile t, assembled in the scratch tile from the
        // tile's pattern entries
#pragma unroll
        for (int e = 0; e < 4; e++) {
          scr[(4 * e + rq) * kTileLd + cq] = 0.0;
          scr[kTilePlane + (4 * e + rq) * kTileLd + cq] = 0.0;
        }
        wave_sync_lds();
        for (int e = p.toff[t] + lane; e < p.toff[t + 1]; e += 64) {
          const double2* g2 = reinterpret_cast<const double2*>(Gw + 10 * (size_t)e);
          double xr = 0.0, xi = 0.0;
#pragma unroll
          for (int nn = 0; nn < 5; nn++) {
            const double2 z = g2[nn];
            xr = fma(er[nn], z.x, fma(-ei[nn], z.y, xr));
            xi = fma(er[nn], z.y, fma(ei[nn], z.x, xi));
          }
          const int pos = p.epos[e];
          scr[(pos >> 4) * kTileLd + (pos & 15)] = xr;
          scr[kTilePlane + (pos >> 4) * kTileLd + (pos & 15)] = xi;
        }
        wave_sync_lds();
        CTile X = tile_load_c(scr, lane);
        // 1-norm (largest column sum of moduli) -> scaling exponent
        double cs = 0.0;
#pragma unroll
        for (int e = 0; e < 4; e++) cs += sqrt(X.r[e] * X.r[e] + X.i[e] * X.i[e]);
        cs += __shfl_xor(cs, 16);
        cs += __shfl_xor(cs, 32);
        double nrm = cs;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) nrm = fmax(nrm, __shfl_xor(nrm, off));
        nrm = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(nrm)),
                               __builtin_amdgcn_readfirstlane(__double2loint(nrm)));  // the same in every lane
        int sexp = 0;
        if (nrm > p.theta) sexp = (int)ceil(log2(nrm / p.theta));
        smax = sexp > smax ? sexp : smax;
        const double scl = ldexp(1.0, -sexp);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          X.r[e] *= scl;
          X.i[e] *= scl;
        }
        // Taylor degree of this tile: the least kt <= K whose remainder bound
        //   sum_{k > kt} x^k / k! <= x^(kt+1) / (kt+1)! / (1 - x / (kt+2))
        // is at most 2^-53 for x = the scaled 1-norm (the plan's K meets it at x = theta = 1)
        const double xs = nrm * scl;
        int kt = K;
        {
          double term = xs;  // x^k / k! at k = 1
          for (int k = 1; k < K; k++) {
            const double next = term * xs / (k + 1);
            if (next <= 1.1102230246251565e-16 * (1.0 - xs / (k + 2))) {
              kt = k;
              break;
            }
            term = next;
          }
        }
        // Paterson-Stockmeyer in blocks of qt <= q: powers X^2 .. X^qt, the top block, then
        // kt/qt Horner steps with Y = X^qt; qt minimises the products (qt - 1) + kt/qt
        int qt = 1;
        for (int c = 2; c <= q; c++)
          if ((c - 1) + kt / c < (qt - 1) + kt / qt) qt = c;
        const int rt = kt / qt;
        double xar[4], xai[4];
        tile_aop(X, scr, lane, xar, xai);
        const CTile X2 = (qt >= 2) ? cmm16(xar, xai, X) : X;
        const CTile X3 = (qt >= 3) ? cmm16(xar, xai, X2) : X2;
        const CTile X4 = (qt >= 4) ? cmm16(xar, xai, X3) : X3;
        const CTile Y = (qt >= 4) ? X4 : ((qt == 3) ? X3 : ((qt == 2) ? X2 : X));
        CTile acc{v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
        tile_taylor_block(acc, rt, qt, kt, sh.tc, X, X2, X3, rq, cq);
        double yar[4], yai[4];
        if (qt == 1) {
#pragma unroll
          for (int s2 = 0; s2 < 4; s2++) {
            yar[s2] = xar[s2];
            yai[s2] = xai[s2];
          }
        } else {
          tile_aop(Y, scr, lane, yar, yai);
        }
        for (int jj = rt - 1; jj >= 0; jj--) {
          acc = cmm16(yar, yai, acc);
          tile_taylor_block(acc, jj, qt, kt, sh.tc, X, X2, X3, rq, cq);
        }
        for (int s = 0; s < sexp; s++) {
          double par[4], pai[4];
          tile_aop(acc, scr, lane, par, pai);
          acc = cmm16(par, pai, acc);
        }
        // U_t <- P U_t
        double par[4], pai[4];
        tile_aop(acc, scr, lane, par, pai);
        const CTile Un = cmm16(par, pai, tile_load_c(sh.u.t.U[t], lane));
        tile_store(sh.u.t.U[t], Un, lane);
        nprod += 4 * ((qt - 1) + rt + sexp + 1);
      }
    }
    __syncthreads();  // U = U_ticks, the period propagator, on every tile
    // rho_k = U rho_{k-1} U' on the pairs: T = U_s rho, rho <- T U_t'
    for (int pp = wave; pp < p.npair; pp += kWaves) {
      const int s = p.pair_s[pp], t = p.pair_t[pp];
      double usr[4], usi[4], utr[4], uti[4];
      tile_load_aop(sh.u.t.U[s], lane, false, usr, usi);  // A operand of U_s
      // B operand of U_t' (C layout of U_t'): element (4e + l/16, l%16) = conj(U_t[l%16][4e + l/16])
      tile_load_aop(sh.u.t.U[t], lane, true, utr, uti);
      const CTile Ut{v4{utr[0], utr[1], utr[2], utr[3]}, v4{uti[0], uti[1], uti[2], uti[3]}};
      CTile rho = tile_global_c(p.rho0_t + (size_t)pp * 512, lane);
      const int32_t* em = p.emap + (size_t)pp * 256;
      for (int k = 0; k < p.n_rho; k++) {
        if (k > 0) {
          const CTile T = cmm16(usr, usi, rho);
          double tar[4], tai[4];
          tile_aop(T, scr, lane, tar, tai);
          rho = cmm16(tar, tai, Ut);
          nprod += 8;
        }
        double* rr_ = Rre + (size_t)k * L;
        double* ri_ = Rim + (size_t)k * L;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int x = em[(4 * e + rq) * kTileN + cq];
          if (x >= 0) {
            rr_[x] = rho.r[e];
            ri_[x] = rho.i[e];
          }
        }
      }
    }
    if (lane == 0) {
      atomicMax(&sh.smax, smax);
      atomicAdd(&sh.nprod, nprod);
    }
    __syncthreads();
    if (tid == 0) {
      if (p.sq_out) p.sq_out[o] = sh.smax;
      if (p.prod_out) p.prod_out[o] = sh.nprod;
    }
    // f(k*ticks + j) = sum conj(C_j) .* rho_k over the detected elements: one product for the
    // rotor periods detected whole, dot products for the ticks of a last, partial period
    {
      const int64_t npts = p.nsteps + 1;
      double* out = p.per_or + 2 * (size_t)o * npts;
      const int ticks = p.ticks;
      const int kfull = (int)(npts / ticks);
      const int rem = (int)(npts - (int64_t)kfull * ticks);
      if (kfull > 0)
        wg_gemm_t<double, false, true, true, false>(ticks, kfull, (int)L, Cre, Cim, (int)L, Rre, Rim, (int)L,
                                                    [=](int jj, int kk, double vr, double vi) {
                                                      const int64_t tt = (int64_t)kk * ticks + jj;
                                                      out[2 * tt] = vr;
                                                      out[2 * tt + 1] = vi;
                                                    },
                                                    sh.u.g);
      if (rem > 0) {
        const double* rr_ = Rre + (size_t)kfull * L;
        const double* ri_ = Rim + (size_t)kfull * L;
        for (int jj = wave; jj < rem; jj += kWaves) {
          const double* cr = Cre + (size_t)jj * L;
          const double* ci = Cim + (size_t)jj * L;
          double sr = 0.0, si = 0.0;
          for (int x = lane; x < L; x += 64) {
            sr = fma(cr[x], rr_[x], fma(ci[x], ri_[x], sr));
            si = fma(cr[x], ri_[x], fma(-ci[x], rr_[x], si));
          }
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off);
            si += __shfl_xor(si, off);
          }
          if (lane == 0) {
            const int64_t tt = (int64_t)kfull * ticks + jj;
            out[2 * tt] = sr;
            out[2 * tt + 1] = si;
          }
        }
      }
    }
  }
}

}  // namespace spk

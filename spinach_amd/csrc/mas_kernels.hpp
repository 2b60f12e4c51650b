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

// Device kernels of the powder propagation library (gfx950, wave64, MFMA).
//
// One work item = one (spin system, orientation) pair. A persistent grid of
// 256-thread workgroups pulls items from a device-side queue. For each item
// the workgroup:
//   1. evaluates the rank-1/rank-2 Wigner matrices of the orientation
//      (kernel/utilities/wigner.m:39-77),
//   2. assembles the dense reduced generator X = -i*dt*(H+H')/2 with
//      H = I + sum_{r,k,m} D^r(k,m) Q{r}{k,m} on the union sparsity pattern
//      (kernel/orientation.m:38-57, kernel/contexts/powder.m:235),
//   3. takes its 1-norm in FP64 (kernel/utilities/cheap_norm.m:50) and picks
//      the scaling exponent s so that |X/2^s|_1 <= theta,
//   4. forms P = expm(X) (the dense branch of kernel/propagator.m:40-42) by a
//      degree-K Taylor polynomial evaluated with the Paterson-Stockmeyer
//      scheme, followed by s squarings,
//   5. evaluates the detected trajectory f(t) = coil' * P^t * rho0,
//      t = 0..nsteps (kernel/evolution.m:562-566) by a two-level power
//      scheme: u_j = coil'*P^j (j < B) built by doubling with the powers
//      P, P^2, .., P^(B/2) that the squarings to Q = P^B produce anyway;
//      v_k = Q^k*rho0 by a matrix-vector chain; f(kB+j) = u_j*v_k as one
//      matrix product U*V^T,
//   6. writes f to the per-orientation buffer.
// A second kernel forms the grid-weighted sum in grid order
// (kernel/contexts/powder.m:280-286).
//
// Real-generator path: when H is real at every orientation (the host checks
// the Q symmetry that guarantees it; secular 'nmr' Hamiltonians in the
// spherical-tensor Liouville basis have it), A = dt*H is real symmetric and
// P = exp(-i*A) = cos(A) - i*sin(A). The Taylor stage then uses real powers
// of A with complex coefficients, and every squaring of P uses the fact that
// its planes commute: two real products instead of four.
//
// All matrix products run on the matrix cores: v_mfma_f64_16x16x4f64 in FP64
// mode, v_mfma_f32_16x16x4f32 in FP32 mode, complex products as four real
// MFMA accumulations. Matrices are planar complex (real plane, then
// imaginary plane), row-major, padded to a multiple of 16 with zero rows and
// columns (the padded block is decoupled and does not change the physical
// block).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

constexpr int kThreads = 256;

// Per-system descriptor, device resident.
constexpr int kRowBlocks = 17;  // row-block boundaries rb[0..16] of a pattern: rows 16k up to 256

struct SysDesc {
  int32_t n;         // reduced dimension
  int32_t np;        // padded dimension (multiple of 16)
  int32_t nnz;       // structural nonzeros of the reduced union pattern
  int32_t nq;        // number of non-empty Q slots
  int64_t pat_off;   // offset into row/col/tpos arrays
  int64_t val_off;   // offset (complex elements) into the value array; (1+nq)*nnz entries
  int64_t slot_off;  // offset into the slot-id array (nq entries)
  int64_t vec_off;   // offset (complex elements) into rho0/coil arrays (n entries each)
  int64_t or_off;    // first orientation of this system in the global orientation arrays
  int64_t n_or;      // number of orientations
  int32_t real;      // 1: the generator is real at every orientation (real-product path)
  int32_t rb[kRowBlocks];  // rb[k]: first pattern entry of row 16k (the pattern is sorted by row, then column)
  int64_t rlx_off;   // relaxation R0 on the union pattern: offset (complex elements) into rvals, -1: none
  int64_t rdiag_off; // relaxation diagonal planes g0..g6 (7*n doubles) in rdiag, -1: none
};

struct Item {
  int32_t sys;
  int32_t pad;
  int64_t orient;  // global orientation index
};

template <typename R>
struct KParams {
  const SysDesc* sys;
  const Item* items;
  int64_t n_items;
  unsigned int* queue;  // work-queue head, zeroed by k_reset before each launch
  const int32_t* prow;
  const int32_t* pcol;
  const int32_t* ptpos;
  const double* vals;    // interleaved complex FP64
  const int32_t* slots;  // Q slot ids 0..33
  const double* rho0;    // interleaved complex
  const double* rho_or;  // optional per-orientation initial states (after a dead time), stride rho_or_stride
  int64_t rho_or_stride;
  const double* coil;    // interleaved complex
  const double* alphas;
  const double* betas;
  const double* gammas;
  double dt;
  int64_t nsteps;
  int32_t block;     // B
  int32_t nblk;      // M = ceil((nsteps+1)/B)
  int32_t log2b;     // log2(B)
  int32_t taylor_k;  // degree K of the Taylor polynomial
  int32_t ps_q;      // Paterson-Stockmeyer block length q
  double theta;      // scaling threshold on the 1-norm
  R* work;           // per-resident-workgroup workspace
  int64_t work_stride;  // R elements per workgroup slot
  int32_t np_max;
  double* per_or;    // per-orientation FIDs: interleaved complex FP64, (nsteps+1) per orientation
  int32_t* sq_out;   // per orientation: scaling exponent used (diagnostics)
  const double* rvals;  // relaxation R0 values (interleaved complex), see SysDesc::rlx_off
  const double* rdiag;  // relaxation diagonal planes, see SysDesc::rdiag_off
};

// q_p(Omega) of the relaxation rate tensors: the quadratic monomials (ox^2, oy^2,
// oz^2, 2 ox oy, 2 ox oz, 2 oy oz) of ort = [0 0 1]*euler2dcm(alpha, beta, gamma)
// = (-sin(beta) cos(gamma), sin(beta) sin(gamma), cos(beta)) (kernel/utilities/
// rlx_t1_t2.m, kernel/relaxation.m 'damp': rate = ort*T*ort')
__host__ __device__ inline void relax_monomials(double beta, double gamma, double q[6]) {
  const double sb = sin(beta), cb = cos(beta), sg = sin(gamma), cg = cos(gamma);
  const double ox = -sb * cg, oy = sb * sg, oz = cb;
  q[0] = ox * ox;
  q[1] = oy * oy;
  q[2] = oz * oz;
  q[3] = 2.0 * ox * oy;
  q[4] = 2.0 * ox * oz;
  q[5] = 2.0 * oy * oz;
}

// zeroes the kQueues work-queue heads of a plan: [0] propagator items with a
// complex generator, [1] tridiagonalisation, [2] synthesis, [3] propagator
// items with a real generator, [4] stepping items with a complex generator,
// [5] stepping items with a real generator, [6] the dead-time step of every item,
// [7] the eigensolver fallback, [8] tridiagonalisation above kTriNP (k_tridiag_w)
constexpr int kQueues = 16;
__global__ void k_reset(unsigned int* q) {
  if (blockIdx.x == 0 && threadIdx.x < kQueues) q[threadIdx.x] = 0u;
}

// ----------------------------------------------------------------- Wigner
// d^l_{m'm}(beta), rows/cols in descending m; closed forms of Brink & Satchler.
__host__ __device__ inline void wigner_small_d2(double b, double d[5][5]) {
  double c = cos(b), s = sin(b);
  double ch = cos(0.5 * b), sh = sin(0.5 * b);
  const double r38 = 0.61237243569579452455, r32 = 1.2247448713915890491;
  double base[3][5];  // m' = 2,1,0 ; m = 2..-2
  base[0][0] = ch * ch * ch * ch;
  base[0][1] = -s * (1 + c) * 0.5;
  base[0][2] = r38 * s * s;
  base[0][3] = -s * (1 - c) * 0.5;
  base[0][4] = sh * sh * sh * sh;
  base[1][1] = (2 * c - 1) * (1 + c) * 0.5;
  base[1][2] = -r32 * s * c;
  base[1][3] = (2 * c + 1) * (1 - c) * 0.5;
  base[2][2] = (3 * c * c - 1) * 0.5;
  const int ms[5] = {2, 1, 0, -1, -2};
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      int mp = ms[i], m = ms[j];
      int a, bb;
      double sg = 1.0;
      if (mp >= (m < 0 ? -m : m)) {
        a = mp; bb = m;
      } else if (-mp >= (m < 0 ? -m : m)) {
        a = -mp; bb = -m; sg = ((mp - m) & 1) ? -1.0 : 1.0;
      } else if (m >= (mp < 0 ? -mp : mp)) {
        a = m; bb = mp; sg = ((m - mp) & 1) ? -1.0 : 1.0;
      } else {
        a = -m; bb = -mp;
      }
      d[i][j] = sg * base[2 - a][2 - bb];
    }
}

__host__ __device__ inline void wigner_small_d1(double b, double d[3][3]) {
  double c = cos(b), s = sin(b);
  const double r2 = 0.70710678118654752440;
  d[0][0] = (1 + c) * 0.5; d[0][1] = -s * r2; d[0][2] = (1 - c) * 0.5;
  d[1][0] = s * r2;        d[1][1] = c;       d[1][2] = -s * r2;
  d[2][0] = (1 - c) * 0.5; d[2][1] = s * r2;  d[2][2] = (1 + c) * 0.5;
}

// coefficient of Q slot t (0..33) at Euler angles (a,b,g): D^r(k,m)
__host__ __device__ inline void wigner_coeffs(double a, double b, double g, double* cre, double* cim) {
  double d1[3][3], d2[5][5];
  wigner_small_d1(b, d1);
  wigner_small_d2(b, d2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double ph = -(1 - i) * a - (1 - j) * g;
      cre[i * 3 + j] = d1[i][j] * cos(ph);
      cim[i * 3 + j] = d1[i][j] * sin(ph);
    }
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double ph = -(2 - i) * a - (2 - j) * g;
      cre[9 + i * 5 + j] = d2[i][j] * cos(ph);
      cim[9 + i * 5 + j] = d2[i][j] * sin(ph);
    }
}

// coefficient of Q slot t (0..33) alone, for one thread per slot: the same
// formulas, so the same values, as wigner_coeffs
// (the element is picked by a select over constant indices, so that the small
// matrices stay in registers)
__host__ __device__ __forceinline__ void wigner_coeff_slot(double a, double b, double g, int t, double* cre, double* cim) {
  double dv = 0.0, ph;
  if (t < 9) {
    double d1[3][3];
    wigner_small_d1(b, d1);
    const int i = t / 3, j = t % 3;
    for (int q = 0; q < 9; q++) dv = (q == t) ? d1[q / 3][q % 3] : dv;
    ph = -(1 - i) * a - (1 - j) * g;
  } else {
    double d2[5][5];
    wigner_small_d2(b, d2);
    const int i = (t - 9) / 5, j = (t - 9) % 5;
    for (int q = 0; q < 25; q++) dv = (q == t - 9) ? d2[q / 5][q % 5] : dv;
    ph = -(2 - i) * a - (2 - j) * g;
  }
  cre[t] = dv * cos(ph);
  cim[t] = dv * sin(ph);
}

// -------------------------------------------------------------- MFMA traits
template <typename R>
struct Mma;

template <>
struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 fma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator element i held by `lane` within a 16x16 tile
  static __device__ inline int row(int lane, int i) { return 4 * i + (lane >> 4); }
};

template <>
struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 fma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

// Planar complex matrix view: element (i,j) at re[i*ld+j], im[i*ld+j]
// (or (j,i) when trans is set); conj negates the imaginary part on load.
template <typename R>
struct CMat {
  const R* re;
  const R* im;
  int ld;
  bool trans;
  bool conj = false;
  __device__ inline size_t off(int i, int j) const {
    return trans ? (size_t)j * ld + i : (size_t)i * ld + j;
  }
};

// LDS staging of one k-panel. Rows are padded by 16 bytes so that the four
// k-rows a wave reads in one MFMA step fall into different banks.
template <typename R>
struct GemmLds {
  static constexpr int LD = 64 + 16 / (int)sizeof(R);
  R ar[16][LD];
  R ai[16][LD];
  R br[16][LD];
  R bi[16][LD];
  __device__ R* flat() { return reinterpret_cast<R*>(this); }  // 4*16*LD >= 4096 elements
};

// Operand modes of wg_gemm_t
constexpr int kCplx = 0;  // complex A, complex B: cr = Ar Br - Ai Bi, ci = Ar Bi + Ai Br
constexpr int kRR = 1;    // real A, real B: cr = Ar Br (imaginary planes are not read)
constexpr int kRC = 2;    // real A, complex B: cr = Ar Br, ci = Ar Bi
constexpr int kSq = 3;    // squaring of P = Pr + i Pi whose planes commute (P a function of one
                          // real symmetric matrix): cr = (Ar - Ai)(Br + Bi) = Re(P^2),
                          // ci = Ar Bi = Im(P^2)/2; the epilogue doubles ci

// C(MxN) = op(A)(MxK) * op(B)(KxN), K a multiple of 16; epi(i, j, re, im)
// receives every output element. TA/TB select transposed storage
// (element (i,k) of A at A[k*lda+i]), CA/CB complex conjugation, MODE the
// operand mode above, all at compile time. The workgroup walks 64x64 output
// tiles; each of its four waves owns a 32x32 quarter as 2x2 MFMA tiles;
// k-panels of 16 are staged through LDS, and the next panel is prefetched
// into registers while the current one feeds the matrix cores.
template <typename R, bool TA, bool CA, bool TB, bool CB, int MODE = kCplx, typename Epi>
__device__ void wg_gemm_t(int M, int N, int K, const R* __restrict__ Are, const R* __restrict__ Aim, int lda,
                          const R* __restrict__ Bre, const R* __restrict__ Bim, int ldb, Epi epi,
                          GemmLds<R>& sm) {
  typedef typename Mma<R>::v4 v4;
  constexpr bool a_cplx = (MODE == kCplx || MODE == kSq);
  constexpr bool b_cplx = (MODE != kRR);
  constexpr bool c_cplx = (MODE != kRR);
  static_assert(MODE != kSq || (!CA && !CB), "the squaring mode takes no conjugation");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  for (int i0 = 0; i0 < M; i0 += 64) {
    for (int j0 = 0; j0 < N; j0 += 64) {
      const bool active = (i0 + 32 * wr < M) && (j0 + 32 * wc < N);
      v4 cr[2][2], ci[2][2];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          cr[a][b] = v4{0, 0, 0, 0};
          ci[a][b] = v4{0, 0, 0, 0};
        }
      // per-thread staging coordinates and panel-0 offsets
      int aoff[4], boff[4];
      bool aok[4], bok[4];
#pragma unroll
      for (int rep = 0; rep < 4; rep++) {
        const int e = tid + rep * kThreads;
        if (!TA) {
          const int r = e >> 4, c = e & 15;
          aok[rep] = (i0 + r < M);
          aoff[rep] = (i0 + r) * lda + c;
        } else {
          const int c = e >> 6, r = e & 63;
          aok[rep] = (i0 + r < M);
          aoff[rep] = c * lda + i0 + r;
        }
        if (!TB) {
          const int rb = e >> 6, cb = e & 63;
          bok[rep] = (j0 + cb < N);
          boff[rep] = rb * ldb + j0 + cb;
        } else {
          const int cb = e >> 4, rb = e & 15;
          bok[rep] = (j0 + cb < N);
          boff[rep] = (j0 + cb) * ldb + rb;
        }
      }
      R pa_r[4], pa_i[4], pb_r[4], pb_i[4];
      auto fetch = [&](int k0) {
        const int da = TA ? k0 * lda : k0;
        const int db = TB ? k0 : k0 * ldb;
#pragma unroll
        for (int rep = 0; rep < 4; rep++) {
          R va = R(0), vb = R(0), wa = R(0), wb = R(0);
          if (aok[rep]) {
            va = Are[aoff[rep] + da];
            if constexpr (a_cplx) {
              vb = Aim[aoff[rep] + da];
              if (CA) vb = -vb;
            }
          }
          if (bok[rep]) {
            wa = Bre[boff[rep] + db];
            if constexpr (b_cplx) {
              wb = Bim[boff[rep] + db];
              if (CB) wb = -wb;
            }
          }
          if constexpr (MODE == kSq) {
            // stage (Ar - Ai, Ar) and (Br + Bi, Bi)
            pa_r[rep] = va - vb;
            pa_i[rep] = va;
            pb_r[rep] = wa + wb;
            pb_i[rep] = wb;
          } else {
            pa_r[rep] = va;
            pa_i[rep] = vb;
            pb_r[rep] = wa;
            pb_i[rep] = wb;
          }
        }
      };
      fetch(0);
      for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int rep = 0; rep < 4; rep++) {
          const int e = tid + rep * kThreads;
          const int r = TA ? (e & 63) : (e >> 4), c = TA ? (e >> 6) : (e & 15);
          sm.ar[c][r] = pa_r[rep];
          if constexpr (a_cplx) sm.ai[c][r] = pa_i[rep];
          const int rb = TB ? (e & 15) : (e >> 6), cb = TB ? (e >> 4) : (e & 63);
          sm.br[rb][cb] = pb_r[rep];
          if constexpr (b_cplx) sm.bi[rb][cb] = pb_i[rep];
        }
        __syncthreads();
        if (k0 + 16 < K) fetch(k0 + 16);
        if (active) {
#pragma unroll
          for (int kk = 0; kk < 16; kk += 4) {
            const int kr = kk + (lane >> 4);
            R xr[2], xi[2], nxi[2], yr[2], yi[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
              xr[t] = sm.ar[kr][32 * wr + 16 * t + (lane & 15)];
              xi[t] = a_cplx ? sm.ai[kr][32 * wr + 16 * t + (lane & 15)] : R(0);
              nxi[t] = -xi[t];
              yr[t] = sm.br[kr][32 * wc + 16 * t + (lane & 15)];
              yi[t] = b_cplx ? sm.bi[kr][32 * wc + 16 * t + (lane & 15)] : R(0);
            }
            if constexpr (MODE == kCplx) {
#pragma unroll
              for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                  cr[a][b] = Mma<R>::fma(xr[a], yr[b], cr[a][b]);
                  ci[a][b] = Mma<R>::fma(xr[a], yi[b], ci[a][b]);
                }
#pragma unroll
              for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                  cr[a][b] = Mma<R>::fma(nxi[a], yi[b], cr[a][b]);
                  ci[a][b] = Mma<R>::fma(xi[a], yr[b], ci[a][b]);
                }
            } else if constexpr (MODE == kRR) {
#pragma unroll
              for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) cr[a][b] = Mma<R>::fma(xr[a], yr[b], cr[a][b]);
            } else if constexpr (MODE == kRC) {
#pragma unroll
              for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                  cr[a][b] = Mma<R>::fma(xr[a], yr[b], cr[a][b]);
                  ci[a][b] = Mma<R>::fma(xr[a], yi[b], ci[a][b]);
                }
            } else {
#pragma unroll
              for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) {
                  cr[a][b] = Mma<R>::fma(xr[a], yr[b], cr[a][b]);
                  ci[a][b] = Mma<R>::fma(xi[a], yi[b], ci[a][b]);
                }
            }
          }
        }
        __syncthreads();
      }
      if (active) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
              int i = i0 + 32 * wr + 16 * a + Mma<R>::row(lane, e);
              int j = j0 + 32 * wc + 16 * b + (lane & 15);
              if (i < M && j < N) epi(i, j, cr[a][b][e], c_cplx ? ci[a][b][e] : R(0));
            }
      }
    }
  }
  __syncthreads();
}

// Convenience form for square planar buffers (no transposition or conjugation).
template <typename R, typename Epi>
__device__ inline void wg_gemm_nn(int M, int N, int K, const R* Are, const R* Aim, int lda, const R* Bre,
                                  const R* Bim, int ldb, Epi epi, GemmLds<R>& sm) {
  wg_gemm_t<R, false, false, false, false>(M, N, K, Are, Aim, lda, Bre, Bim, ldb, epi, sm);
}

// ---------------------------------------------------------- block reductions
__device__ inline double wg_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __syncthreads();
  return r;
}

template <typename R>
struct SharedAll {
  GemmLds<R> g;
  R vpart[4 * 64 * 2];  // per-wave partial sums of the matrix-vector chain
  double cre[34], cim[34];
  double red[4];
  double tc[32];  // Taylor coefficients 1/k!
  int item;
  int sexp;
};

// Planar complex square matrix buffer of the workspace
template <typename R>
struct Buf {
  R* p;
  size_t pl;  // plane size np*np
  int np;
  __device__ inline R* re() const { return p; }
  __device__ inline R* im() const { return p + pl; }
  __device__ inline CMat<R> view(bool trans = false, bool cj = false) const {
    return CMat<R>{p, p + pl, np, trans, cj};
  }
};

// P = expm(X) for a pre-scaled X (|X|_1 <= theta) held in bf[0], followed by
// sexp squarings. The Taylor polynomial of degree K is evaluated by the
// Paterson-Stockmeyer scheme: powers X^2..X^q go to bf[1..q-1], then Horner
// in Y = X^q over blocks B_j(X) = sum_{i<q} c_{jq+i} X^i (c_k = tc[k] = 1/k!).
// Returns the buffer that holds P; every other buffer of bf may be reused.
template <typename R>
__device__ Buf<R> wg_expm(Buf<R>* bf, int sexp, int q, int K, const double* tc, GemmLds<R>& sm) {
  const int tid = threadIdx.x;
  const int np = bf[0].np;
  const size_t pl = bf[0].pl;
  for (int k = 2; k <= q; k++) {
    R* dre = bf[k - 1].re();
    R* dim_ = bf[k - 1].im();
    wg_gemm_nn<R>(np, np, np, bf[k - 2].re(), bf[k - 2].im(), np, bf[0].re(), bf[0].im(), np,
               [=](int i, int j, R vr, R vi) {
                 size_t o2 = (size_t)i * np + j;
                 dre[o2] = vr;
                 dim_[o2] = vi;
               },
               sm);
  }
  const int r = K / q;
  Buf<R> acc = bf[4], tmp = bf[5];
  const Buf<R> b1 = bf[0], b2 = bf[1], b3 = bf[2];
  // top block: degrees r*q .. K
  for (size_t e = tid; e < pl; e += kThreads) {
    double ar = 0.0, ai = 0.0;
    for (int i2 = 0; i2 < q; i2++) {
      int k = r * q + i2;
      if (k > K) break;
      if (i2 == 0) {
        if ((e / np) == (e % np)) ar += tc[k];
      } else {
        const Buf<R>& bb = (i2 == 1) ? b1 : ((i2 == 2) ? b2 : b3);
        ar += tc[k] * (double)bb.re()[e];
        ai += tc[k] * (double)bb.im()[e];
      }
    }
    acc.re()[e] = (R)ar;
    acc.im()[e] = (R)ai;
  }
  __syncthreads();
  const Buf<R> Y = bf[q - 1];
  for (int j = r - 1; j >= 0; j--) {
    R* dre = tmp.re();
    R* dim_ = tmp.im();
    wg_gemm_nn<R>(np, np, np, Y.re(), Y.im(), np, acc.re(), acc.im(), np,
               [=](int i, int jj, R vr, R vi) {
                 size_t o2 = (size_t)i * np + jj;
                 double ar = (double)vr, ai = (double)vi;
                 for (int i2 = 0; i2 < q; i2++) {
                   int k = j * q + i2;
                   if (k > K) break;
                   if (i2 == 0) {
                     if (i == jj) ar += tc[k];
                   } else {
                     const Buf<R>& bb = (i2 == 1) ? b1 : ((i2 == 2) ? b2 : b3);
                     ar += tc[k] * (double)bb.re()[o2];
                     ai += tc[k] * (double)bb.im()[o2];
                   }
                 }
                 dre[o2] = (R)ar;
                 dim_[o2] = (R)ai;
               },
               sm);
    Buf<R> t2 = acc;
    acc = tmp;
    tmp = t2;
  }
  // squarings, ping-pong through X and X^2 (no longer needed)
  Buf<R> cur = acc;
  Buf<R> spare[2] = {bf[0], bf[1]};
  int sp_i = 0;
  for (int s = 0; s < sexp; s++) {
    Buf<R> dst = spare[sp_i];
    R* dre = dst.re();
    R* dim_ = dst.im();
    wg_gemm_nn<R>(np, np, np, cur.re(), cur.im(), np, cur.re(), cur.im(), np,
               [=](int i, int j, R vr, R vi) {
                 size_t o2 = (size_t)i * np + j;
                 dre[o2] = vr;
                 dim_[o2] = vi;
               },
               sm);
    spare[sp_i] = cur;
    cur = dst;
    sp_i ^= 1;
  }
  return cur;
}

// Squaring dst = src^2 of a propagator whose real and imaginary planes
// commute (two real products instead of four); `trans` stores the result
// transposed.
template <typename R>
__device__ inline void wg_square_commuting(const Buf<R>& src, const Buf<R>& dst, bool trans, GemmLds<R>& sm) {
  const int np = src.np;
  R* dre = dst.re();
  R* dim_ = dst.im();
  wg_gemm_t<R, false, false, false, false, kSq>(np, np, np, src.re(), src.im(), np, src.re(), src.im(), np,
                                                [=](int i, int j, R vr, R vi) {
                                                  size_t o2 = trans ? (size_t)j * np + i : (size_t)i * np + j;
                                                  dre[o2] = vr;
                                                  dim_[o2] = vi + vi;
                                                },
                                                sm);
}

// P = exp(-i*A) for a real symmetric, pre-scaled A (|A|_1 <= theta) held in
// the real plane of bf[0], followed by sexp squarings. The Taylor polynomial
// sum_k c_k A^k with c_k = (-i)^k/k! has real powers and complex
// coefficients: powers A^2..A^q are real products into the real planes of
// bf[1..q-1]; each Horner step Y*acc (Y = A^q real, acc complex) is two real
// products; each squaring of P = cos(A) - i sin(A) is two real products,
// Re(P^2) = (C+S)(C-S) and Im(P^2) = -2CS. Returns the buffer that holds P.
template <typename R>
__device__ Buf<R> wg_expm_real(Buf<R>* bf, int sexp, int q, int K, const double* tc, GemmLds<R>& sm) {
  const int tid = threadIdx.x;
  const int np = bf[0].np;
  const size_t pl = bf[0].pl;
  for (int k = 2; k <= q; k++) {
    R* dre = bf[k - 1].re();
    wg_gemm_t<R, false, false, false, false, kRR>(np, np, np, bf[k - 2].re(), nullptr, np, bf[0].re(), nullptr, np,
                                                  [=](int i, int j, R vr, R) { dre[(size_t)i * np + j] = vr; }, sm);
  }
  // c_k = (-i)^k/k!: real part +tc[k] (k%4==0), -tc[k] (k%4==2); imaginary -tc[k] (k%4==1), +tc[k] (k%4==3)
  const R* p1 = bf[0].re();
  const R* p2 = bf[1].re();
  const R* p3 = bf[2].re();
  auto add_block = [=](int jb, size_t e, bool diag, double& ar, double& ai) {
    for (int i2 = 0; i2 < q; i2++) {
      const int k = jb * q + i2;
      if (k > K) break;
      const double c = ((k & 2) ? -tc[k] : tc[k]);
      const R* pw = (i2 == 1) ? p1 : ((i2 == 2) ? p2 : p3);
      const double v = (i2 == 0) ? (diag ? 1.0 : 0.0) : (double)pw[e];
      if (k & 1)
        ai -= c * v;
      else
        ar += c * v;
    }
  };
  const int r = K / q;
  Buf<R> acc = bf[4], tmp = bf[5];
  for (size_t e = tid; e < pl; e += kThreads) {
    double ar = 0.0, ai = 0.0;
    add_block(r, e, (e / np) == (e % np), ar, ai);
    acc.re()[e] = (R)ar;
    acc.im()[e] = (R)ai;
  }
  __syncthreads();
  const R* Y = bf[q - 1].re();
  for (int j = r - 1; j >= 0; j--) {
    R* dre = tmp.re();
    R* dim_ = tmp.im();
    wg_gemm_t<R, false, false, false, false, kRC>(np, np, np, Y, nullptr, np, acc.re(), acc.im(), np,
                                                  [=](int i, int jj, R vr, R vi) {
                                                    size_t o2 = (size_t)i * np + jj;
                                                    double ar = (double)vr, ai = (double)vi;
                                                    add_block(j, o2, i == jj, ar, ai);
                                                    dre[o2] = (R)ar;
                                                    dim_[o2] = (R)ai;
                                                  },
                                                  sm);
    Buf<R> t2 = acc;
    acc = tmp;
    tmp = t2;
  }
  Buf<R> cur = acc;
  Buf<R> spare[2] = {bf[0], bf[1]};
  int sp_i = 0;
  for (int s = 0; s < sexp; s++) {
    Buf<R> dst = spare[sp_i];
    wg_square_commuting<R>(cur, dst, false, sm);
    spare[sp_i] = cur;
    cur = dst;
    sp_i ^= 1;
  }
  return cur;
}

// --------------------------------------------------------- the fused kernel
// REAL: the real-generator instantiation (real-product expm and commuting
// squarings); each instantiation carries one variant, so the host launches the
// real and the complex items of a plan separately.
template <typename R, bool REAL>
__global__ void __launch_bounds__(kThreads, 2) k_powder(KParams<R> p) {
  __shared__ SharedAll<R> sh;
  const int tid = threadIdx.x;
  R* ws = p.work + (size_t)blockIdx.x * p.work_stride;
  const size_t mat = 2 * (size_t)p.np_max * p.np_max;
  R* bufbase[6];
  for (int b = 0; b < 6; b++) bufbase[b] = ws + b * mat;
  R* Ubase = ws + 6 * mat;  // (B + 16) rows x np, planar
  R* Vbase = Ubase + 2 * (size_t)(p.block + 16) * p.np_max;  // nblk rows x np, planar

  if (tid < 32) {
    double f = 1.0;
    for (int k = 2; k <= tid; k++) f *= k;
    sh.tc[tid] = 1.0 / f;
  }

  for (;;) {
    if (tid == 0) sh.item = (int)atomicAdd(p.queue, 1u);
    __syncthreads();
    const int64_t it = sh.item;
    __syncthreads();
    if (it >= p.n_items) break;
    const Item item = p.items[it];
    const SysDesc sd = p.sys[item.sys];
    const int np = sd.np;
    const size_t pl = (size_t)np * np;
    const int64_t o = item.orient;
    Buf<R> bf[6];
    for (int b = 0; b < 6; b++) bf[b] = Buf<R>{bufbase[b], pl, np};

    // 1. Wigner coefficients
    if (tid == 0) wigner_coeffs(p.alphas[o], p.betas[o], p.gammas[o], sh.cre, sh.cim);
    // 2. zero X and the U block
    R* X = bf[0].p;
    for (size_t e = tid; e < 2 * pl; e += kThreads) X[e] = R(0);
    const size_t ub = (size_t)(p.block + 16) * np;  // plane size of U
    for (size_t e = tid; e < 2 * ub; e += kThreads) Ubase[e] = R(0);
    __syncthreads();
    // scatter the symmetrised generator: X = -i*dt*(x + x^H)/2
    {
      const int32_t* prow = p.prow + sd.pat_off;
      const int32_t* pcol = p.pcol + sd.pat_off;
      const int32_t* ptp = p.ptpos + sd.pat_off;
      const double* vals = p.vals + 2 * sd.val_off;
      const int32_t* slots = p.slots + sd.slot_off;
      const int nnz = sd.nnz;
      for (int e = tid; e < nnz; e += kThreads) {
        int et = ptp[e];
        double xr = vals[2 * e], xi = vals[2 * e + 1];
        double tr = vals[2 * et], ti = vals[2 * et + 1];
        for (int q = 0; q < sd.nq; q++) {
          const double* qv = vals + 2 * (size_t)(q + 1) * nnz;
          double cr = sh.cre[slots[q]], ci = sh.cim[slots[q]];
          double vr = qv[2 * e], vi = qv[2 * e + 1];
          xr += cr * vr - ci * vi;
          xi += cr * vi + ci * vr;
          double wr = qv[2 * et], wi = qv[2 * et + 1];
          tr += cr * wr - ci * wi;
          ti += cr * wi + ci * wr;
        }
        double hr = 0.5 * (xr + tr), hi = 0.5 * (xi - ti);
        size_t off = (size_t)prow[e] * np + pcol[e];
        if constexpr (REAL) {
          X[off] = (R)(p.dt * hr);  // A = dt*H; the imaginary parts cancel pairwise
        } else {
          // L = H + i R: X = -i*dt*L = -i*dt*H + dt*R (R is not symmetrised)
          double rr = 0.0, ri = 0.0;
          if (sd.rlx_off >= 0) {
            rr = p.rvals[2 * (sd.rlx_off + e)];
            ri = p.rvals[2 * (sd.rlx_off + e) + 1];
          }
          X[off] = (R)(p.dt * (hi + rr));
          X[pl + off] = (R)(p.dt * (ri - hr));
        }
      }
    }
    __syncthreads();
    if constexpr (!REAL) {
      // the diagonal relaxation terms of this orientation: R_ii += g0 + sum_p q_p g_p
      if (sd.rdiag_off >= 0) {
        double q[6];
        relax_monomials(p.betas[o], p.gammas[o], q);
        const double* g = p.rdiag + sd.rdiag_off;
        for (int i = tid; i < sd.n; i += kThreads) {
          double r = g[i];
          for (int k = 0; k < 6; k++) r = fma(q[k], g[(size_t)(k + 1) * sd.n + i], r);
          const size_t off = (size_t)i * np + i;
          X[off] = (R)((double)X[off] + p.dt * r);
        }
        __syncthreads();
      }
    }
    // 3. FP64 1-norm (max column sum) and scaling exponent
    {
      double colmax = 0.0;
      for (int j = tid; j < np; j += kThreads) {
        double s = 0.0;
        for (int i = 0; i < np; i++) {
          double a = (double)X[(size_t)i * np + j], b = (double)X[pl + (size_t)i * np + j];
          s += sqrt(a * a + b * b);
        }
        colmax = fmax(colmax, s);
      }
      double nrm = wg_max(colmax, sh.red);
      if (tid == 0) {
        int s = 0;
        if (nrm > p.theta) s = (int)ceil(log2(nrm / p.theta));
        sh.sexp = s;
        if (p.sq_out) p.sq_out[o] = s;
      }
      __syncthreads();
    }
    const int sexp = sh.sexp;
    {
      const R scl = (R)ldexp(1.0, -sexp);
      for (size_t e = tid; e < 2 * pl; e += kThreads) X[e] *= scl;
      __syncthreads();
    }
    // 4. P = expm(X): Paterson-Stockmeyer Taylor polynomial and s squarings
    Buf<R> cur;
    if constexpr (REAL)
      cur = wg_expm_real<R>(bf, sexp, p.ps_q, p.taylor_k, sh.tc, sh.g);
    else
      cur = wg_expm<R>(bf, sexp, p.ps_q, p.taylor_k, sh.tc, sh.g);
    Buf<R> spare[2];
    {
      int ns = 0;
      for (int b2 = 0; b2 < 6 && ns < 2; b2++)
        if (b2 != 2 && bf[b2].p != cur.p) spare[ns++] = bf[b2];
    }
    int sp_i = 0;
    // cur = P. 5a. U doubling with P^(2^k) and squarings up to Q = P^B (stored transposed in QT)
    const double* coilv = p.coil + 2 * sd.vec_off;
    const double* rhov = p.rho_or ? p.rho_or + 2 * (size_t)o * p.rho_or_stride : p.rho0 + 2 * sd.vec_off;
    R* Ure = Ubase;
    R* Uim = Ubase + ub;
    const size_t vb = (size_t)p.nblk * np;
    R* Vre = Vbase;
    R* Vim = Vbase + vb;
    for (int c = tid; c < np; c += kThreads) {
      R ur = R(0), ui = R(0), vr = R(0), vi = R(0);
      if (c < sd.n) {
        ur = (R)coilv[2 * c];
        ui = (R)(-coilv[2 * c + 1]);
        vr = (R)rhov[2 * c];
        vi = (R)rhov[2 * c + 1];
      }
      Ure[c] = ur;
      Uim[c] = ui;
      Vre[c] = vr;
      Vim[c] = vi;
    }
    __syncthreads();
    Buf<R> QT = bf[2];  // free after the Taylor stage
    if (cur.p == QT.p) QT = bf[3];
    for (int k = 0; k < p.log2b; k++) {
      const int rows = 1 << k;
      // U[rows .. 2*rows) = U[0 .. rows) * P^(2^k)
      {
        R* ure2 = Ure + (size_t)rows * np;
        R* uim2 = Uim + (size_t)rows * np;
        wg_gemm_nn<R>(rows, np, np, Ure, Uim, np, cur.re(), cur.im(), np,
                   [=](int i, int j, R vr, R vi) {
                     ure2[(size_t)i * np + j] = vr;
                     uim2[(size_t)i * np + j] = vi;
                   },
                   sh.g);
      }
      // next power; the last one is Q = P^B, stored transposed
      const bool last = (k == p.log2b - 1);
      Buf<R> dst = last ? QT : spare[sp_i];
      if (!last && dst.p == QT.p) dst = spare[sp_i ^ 1];
      if constexpr (REAL) {
        wg_square_commuting<R>(cur, dst, last, sh.g);
      } else {
        R* dre = dst.re();
        R* dim_ = dst.im();
        wg_gemm_nn<R>(np, np, np, cur.re(), cur.im(), np, cur.re(), cur.im(), np,
                   [=](int i, int j, R vr, R vi) {
                     size_t o2 = last ? (size_t)j * np + i : (size_t)i * np + j;
                     dre[o2] = vr;
                     dim_[o2] = vi;
                   },
                   sh.g);
      }
      if (!last) {
        spare[sp_i] = cur;
        cur = dst;
        sp_i ^= 1;
      }
    }
    if (p.log2b == 0) {
      for (size_t e = tid; e < pl; e += kThreads) {
        size_t i = e / np, j = e % np;
        QT.re()[j * np + i] = cur.re()[e];
        QT.im()[j * np + i] = cur.im()[e];
      }
      __syncthreads();
    }
    // 5b. V chain: v_{k+1}[r] = sum_c Q[r][c] v_k[c] = sum_c QT[c][r] v_k[c].
    //     The vector lives in LDS (the GEMM staging area is idle here); each
    //     wave sums a quarter of the columns for 64 rows at a time with
    //     coalesced loads of QT, and the four partial sums meet in LDS.
    {
      const R* qre = QT.re();
      const R* qim = QT.im();
      R* va = sh.g.flat();  // 4*np <= 4096 elements: two ping-pong vectors
      R* vb = va + 2 * np;
      const int lane = tid & 63, wave = tid >> 6;
      const int cq = np / 4;
      const int cb0 = wave * cq;
      for (int c = tid; c < np; c += kThreads) {
        va[c] = Vre[c];
        va[np + c] = Vim[c];
      }
      __syncthreads();
      for (int k = 1; k < p.nblk; k++) {
        for (int r0 = 0; r0 < np; r0 += 64) {
          const int rr = r0 + lane;
          R ar = R(0), ai = R(0);
          if (rr < np) {
            const R* q1 = qre + (size_t)cb0 * np + rr;
            const R* q2 = qim + (size_t)cb0 * np + rr;
#pragma unroll 8
            for (int c = 0; c < cq; c++) {
              const R a = q1[(size_t)c * np], b = q2[(size_t)c * np];
              const R xr = va[cb0 + c], xi = va[np + cb0 + c];
              ar = fma(a, xr, ar);
              ar = fma(-b, xi, ar);
              ai = fma(a, xi, ai);
              ai = fma(b, xr, ai);
            }
          }
          sh.vpart[2 * (wave * 64 + lane)] = ar;
          sh.vpart[2 * (wave * 64 + lane) + 1] = ai;
          __syncthreads();
          if (tid < 64 && r0 + tid < np) {
            R sr = R(0), si = R(0);
#pragma unroll
            for (int w = 0; w < 4; w++) {
              sr += sh.vpart[2 * (w * 64 + tid)];
              si += sh.vpart[2 * (w * 64 + tid) + 1];
            }
            vb[r0 + tid] = sr;
            vb[np + r0 + tid] = si;
            Vre[(size_t)k * np + r0 + tid] = sr;
            Vim[(size_t)k * np + r0 + tid] = si;
          }
          __syncthreads();
        }
        R* t = va;
        va = vb;
        vb = t;
      }
      __syncthreads();
    }
    // 5c. F = U * V^T ; f(kB + j) = F[j][k]
    {
      const int64_t npts = p.nsteps + 1;
      double* out = p.per_or + 2 * (size_t)o * npts;
      const int B = p.block;
      wg_gemm_t<R, false, false, true, false>(B, p.nblk, np, Ure, Uim, np, Vre, Vim, np,
                 [=](int j, int k, R vr, R vi) {
                   int64_t t = (int64_t)k * B + j;
                   if (t < npts) {
                     out[2 * t] = (double)vr;
                     out[2 * t + 1] = (double)vi;
                   }
                 },
                 sh.g);
    }
  }
}

// Grid-weighted orientation sum (kernel/contexts/powder.m:280-286) in two
// fixed-order FP64 passes that fill the GPU: chunks of up to kSumChunk
// consecutive orientations of one system (one grid row per chunk, one thread
// per time point), then the chunks of each system in order.
constexpr int kSumChunk = 64;

struct SumChunk {
  int64_t first;  // global index of the chunk's first orientation
  int32_t count;
  int32_t pad;
};

__global__ void k_chunk_sum(const double* __restrict__ per_or, const double* __restrict__ weights,
                            const SumChunk* __restrict__ chunks, int64_t npts, double* __restrict__ partial) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const SumChunk ch = chunks[blockIdx.y];
  double sr = 0.0, si = 0.0;
  for (int k = 0; k < ch.count; k++) {
    const int64_t o = ch.first + k;
    const double w = weights[o];
    const double* f = per_or + 2 * ((size_t)o * npts + t);
    sr = fma(w, f[0], sr);
    si = fma(w, f[1], si);
  }
  double* out = partial + 2 * ((size_t)blockIdx.y * npts + t);
  out[0] = sr;
  out[1] = si;
}

__global__ void k_system_sum(const double* __restrict__ partial, const int64_t* __restrict__ chunk_off, int64_t npts,
                             double* __restrict__ fid) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const int64_t s = blockIdx.y;
  double sr = 0.0, si = 0.0;
  for (int64_t c = chunk_off[s]; c < chunk_off[s + 1]; c++) {
    const double* f = partial + 2 * ((size_t)c * npts + t);
    sr += f[0];
    si += f[1];
  }
  fid[2 * ((size_t)s * npts + t)] = sr;
  fid[2 * ((size_t)s * npts + t) + 1] = si;
}

// Zero fill (a kernel, not a memset node, so that a captured pass replays it).
__global__ void k_fill_zero(double* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// fid += partial[0] + partial[1] + ... in chunk order, one addition at a time:
// block after block this repeats k_system_sum's sequence of additions.
__global__ void k_accumulate(const double* __restrict__ partial, int32_t n_chunks, int64_t npts,
                             double* __restrict__ fid) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  double sr = fid[2 * t], si = fid[2 * t + 1];
  for (int32_t c = 0; c < n_chunks; c++) {
    const double* f = partial + 2 * ((size_t)c * npts + t);
    sr += f[0];
    si += f[1];
  }
  fid[2 * t] = sr;
  fid[2 * t + 1] = si;
}

}  // namespace spk

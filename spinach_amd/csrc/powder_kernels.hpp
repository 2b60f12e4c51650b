// Device kernels of the powder propagation library (gfx950, wave64).
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
//      the scaling exponent s so that |X/2^s| <= theta,
//   4. forms P = expm(X) by a degree-K Taylor polynomial in Horner form and s
//      squarings (the dense branch of kernel/propagator.m:40-42),
//   5. evaluates the detected trajectory f(t) = coil' * P^t * rho0,
//      t = 0..nsteps (kernel/evolution.m:562-566), by a two-level power
//      scheme: Q = P^B, u_j = coil'*P^j (j<B), v_k = Q^k*rho0, f(kB+j) = u_j*v_k,
//   6. writes f to the per-orientation buffer.
// A second kernel forms the grid-weighted sum in grid order
// (kernel/contexts/powder.m:280-286).
//
// Matrices are planar complex (real plane, then imaginary plane), row-major,
// padded to a multiple of 16 with zero rows/columns (the padded block is
// decoupled and does not change the physical block).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

constexpr int kThreads = 256;

// Per-system descriptor, device resident.
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
  const double* vals;   // interleaved complex FP64
  const int32_t* slots;  // Q slot ids 0..33
  const double* rho0;   // interleaved complex
  const double* coil;   // interleaved complex
  const double* alphas;
  const double* betas;
  const double* gammas;
  double dt;
  int64_t nsteps;
  int32_t block;     // B
  int32_t nblk;      // M = ceil((nsteps+1)/B)
  int32_t log2b;     // log2(B)
  int32_t taylor_k;  // K
  double theta;      // scaling threshold
  R* work;           // per-resident-workgroup workspace
  int64_t work_stride;  // R elements per workgroup slot
  int32_t np_max;
  double* per_or;    // per-orientation FIDs: interleaved complex FP64, (nsteps+1) per orientation
  int32_t* sq_out;   // per orientation: scaling exponent used (diagnostics)
};

__global__ void k_reset(unsigned int* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0u;
}

// ----------------------------------------------------------------- Wigner
// d^l_{m'm}(beta), rows/cols in descending m; closed forms of Brink & Satchler.
__device__ inline void wigner_small_d2(double b, double d[5][5]) {
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

__device__ inline void wigner_small_d1(double b, double d[3][3]) {
  double c = cos(b), s = sin(b);
  const double r2 = 0.70710678118654752440;
  d[0][0] = (1 + c) * 0.5; d[0][1] = -s * r2; d[0][2] = (1 - c) * 0.5;
  d[1][0] = s * r2;        d[1][1] = c;       d[1][2] = -s * r2;
  d[2][0] = (1 - c) * 0.5; d[2][1] = s * r2;  d[2][2] = (1 + c) * 0.5;
}

// coefficient of Q slot t (0..33) at Euler angles (a,b,g): D^r(k,m)
__device__ inline void wigner_coeffs(double a, double b, double g, double* cre, double* cim) {
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

// ------------------------------------------------------------ workgroup GEMM
// C = alpha*A*B (+ I if add_identity); optional transposed store of C.
// 64x64 output tiles, k-panels of 16 staged in LDS, 4x4 micro-tile per thread.
template <typename R>
struct GemmLds {
  R ar[16][64];
  R ai[16][64];
  R br[16][64];
  R bi[16][64];
};

template <typename R>
__device__ void wg_zgemm(const R* __restrict__ A, const R* __restrict__ B, R* __restrict__ C, int np,
                         R alpha, bool add_identity, bool transpose_out, GemmLds<R>& sm) {
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const size_t pl = (size_t)np * np;
  for (int i0 = 0; i0 < np; i0 += 64) {
    for (int j0 = 0; j0 < np; j0 += 64) {
      R cr[4][4], ci[4][4];
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++) {
          cr[m][n] = R(0);
          ci[m][n] = R(0);
        }
      for (int k0 = 0; k0 < np; k0 += 16) {
#pragma unroll
        for (int e = tid; e < 1024; e += kThreads) {
          int r = e >> 4, c = e & 15;
          int gi = i0 + r;
          R va = R(0), vb = R(0);
          if (gi < np) {
            size_t o = (size_t)gi * np + k0 + c;
            va = A[o];
            vb = A[pl + o];
          }
          sm.ar[c][r] = va;
          sm.ai[c][r] = vb;
          int rb = e >> 6, cb = e & 63;
          int gj = j0 + cb;
          R wa = R(0), wb = R(0);
          if (gj < np) {
            size_t o = (size_t)(k0 + rb) * np + gj;
            wa = B[o];
            wb = B[pl + o];
          }
          sm.br[rb][cb] = wa;
          sm.bi[rb][cb] = wb;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; kk++) {
          R xr[4], xi[4], yr[4], yi[4];
#pragma unroll
          for (int m = 0; m < 4; m++) {
            xr[m] = sm.ar[kk][ty + 16 * m];
            xi[m] = sm.ai[kk][ty + 16 * m];
            yr[m] = sm.br[kk][tx + 16 * m];
            yi[m] = sm.bi[kk][tx + 16 * m];
          }
#pragma unroll
          for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 4; n++) {
              cr[m][n] = fma(xr[m], yr[n], cr[m][n]);
              cr[m][n] = fma(-xi[m], yi[n], cr[m][n]);
              ci[m][n] = fma(xr[m], yi[n], ci[m][n]);
              ci[m][n] = fma(xi[m], yr[n], ci[m][n]);
            }
        }
        __syncthreads();
      }
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++) {
          int i = i0 + ty + 16 * m, j = j0 + tx + 16 * n;
          if (i < np && j < np) {
            R vr = alpha * cr[m][n], vi = alpha * ci[m][n];
            if (add_identity && i == j) vr += R(1);
            size_t o = transpose_out ? (size_t)j * np + i : (size_t)i * np + j;
            C[o] = vr;
            C[pl + o] = vi;
          }
        }
    }
  }
  __syncthreads();
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
  double cre[34], cim[34];
  double red[4];
  int item;
  int sexp;
};

// --------------------------------------------------------- the fused kernel
template <typename R>
__global__ void __launch_bounds__(kThreads) k_powder(KParams<R> p) {
  __shared__ SharedAll<R> sh;
  const int tid = threadIdx.x;
  R* ws = p.work + (size_t)blockIdx.x * p.work_stride;
  const size_t mat = 2 * (size_t)p.np_max * p.np_max;
  R* X = ws;
  R* T0 = ws + mat;
  R* T1 = ws + 2 * mat;
  R* P = ws + 3 * mat;
  R* QT = ws + 4 * mat;  // Q = P^B stored transposed
  R* U = ws + 5 * mat;   // B rows x np (planar: re block, im block)
  R* V = U + 2 * (size_t)p.block * p.np_max;  // M rows x np

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

    // 1. Wigner coefficients
    if (tid == 0) wigner_coeffs(p.alphas[o], p.betas[o], p.gammas[o], sh.cre, sh.cim);
    // 2. zero X
    for (size_t e = tid; e < 2 * pl; e += kThreads) X[e] = R(0);
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
        // h = (x + conj(x_t))/2 ; X = -i*dt*h
        double hr = 0.5 * (xr + tr), hi = 0.5 * (xi - ti);
        size_t off = (size_t)prow[e] * np + pcol[e];
        X[off] = (R)(p.dt * hi);
        X[pl + off] = (R)(-p.dt * hr);
      }
    }
    __syncthreads();
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
    const R scl = (R)ldexp(1.0, -sexp);
    for (size_t e = tid; e < 2 * pl; e += kThreads) X[e] *= scl;
    // 4. Taylor in Horner form: T = I + X/K ; T = I + X*T/k for k=K-1..1
    {
      const R invk = (R)(1.0 / p.taylor_k);
      for (size_t e = tid; e < 2 * pl; e += kThreads) {
        size_t rem = e % pl;
        R v = X[e] * invk;
        if (e < pl && (rem / np) == (rem % np)) v += R(1);
        T0[e] = v;
      }
      __syncthreads();
    }
    R* cur = T0;
    R* nxt = T1;
    for (int k = p.taylor_k - 1; k >= 1; k--) {
      wg_zgemm<R>(X, cur, nxt, np, (R)(1.0 / k), true, false, sh.g);
      R* t = cur; cur = nxt; nxt = t;
    }
    // squarings
    for (int s = 0; s < sexp; s++) {
      wg_zgemm<R>(cur, cur, nxt, np, R(1), false, false, sh.g);
      R* t = cur; cur = nxt; nxt = t;
    }
    // P <- cur
    for (size_t e = tid; e < 2 * pl; e += kThreads) P[e] = cur[e];
    __syncthreads();
    // 5a. Q = P^B by log2(B) squarings; last one stored transposed
    {
      const R* src = P;
      R* bufs[2] = {T0, T1};
      if (p.log2b == 0) {
        for (size_t e = tid; e < 2 * pl; e += kThreads) {
          size_t rem = e % pl, i = rem / np, j = rem % np;
          QT[(e < pl ? 0 : pl) + j * np + i] = P[e];
        }
        __syncthreads();
      }
      for (int s = 0; s < p.log2b; s++) {
        bool last = (s == p.log2b - 1);
        R* dst = last ? QT : bufs[s & 1];
        wg_zgemm<R>(src, src, dst, np, R(1), false, last, sh.g);
        src = dst;
      }
    }
    // 5b. u_0 = conj(coil) (row), u_{j+1} = u_j * P ; v_0 = rho0, v_{k+1} = Q v_k
    {
      const double* coil = p.coil + 2 * sd.vec_off;
      const double* rho = p.rho0 + 2 * sd.vec_off;
      const size_t ub = (size_t)p.block * np;   // plane size of U
      const size_t vb = (size_t)p.nblk * np;    // plane size of V
      for (int c = tid; c < np; c += kThreads) {
        R ur = R(0), ui = R(0), vr = R(0), vi = R(0);
        if (c < sd.n) {
          ur = (R)coil[2 * c];
          ui = (R)(-coil[2 * c + 1]);
          vr = (R)rho[2 * c];
          vi = (R)rho[2 * c + 1];
        }
        U[c] = ur; U[ub + c] = ui;
        V[c] = vr; V[vb + c] = vi;
      }
      __syncthreads();
      // U chain: u_{j+1}[c] = sum_r u_j[r] P[r][c]
      for (int j = 1; j < p.block; j++) {
        const R* ur = U + (size_t)(j - 1) * np;
        const R* ui = U + ub + (size_t)(j - 1) * np;
        for (int c = tid; c < np; c += kThreads) {
          R ar = R(0), ai = R(0);
          for (int r = 0; r < np; r++) {
            R pr = P[(size_t)r * np + c], pi = P[pl + (size_t)r * np + c];
            R xr = ur[r], xi = ui[r];
            ar = fma(xr, pr, ar); ar = fma(-xi, pi, ar);
            ai = fma(xr, pi, ai); ai = fma(xi, pr, ai);
          }
          U[(size_t)j * np + c] = ar;
          U[ub + (size_t)j * np + c] = ai;
        }
        __syncthreads();
      }
      // V chain: v_{k+1}[r] = sum_c Q[r][c] v_k[c] = sum_c QT[c][r] v_k[c]
      for (int k = 1; k < p.nblk; k++) {
        const R* xr_ = V + (size_t)(k - 1) * np;
        const R* xi_ = V + vb + (size_t)(k - 1) * np;
        for (int r = tid; r < np; r += kThreads) {
          R ar = R(0), ai = R(0);
          for (int c = 0; c < np; c++) {
            R qr = QT[(size_t)c * np + r], qi = QT[pl + (size_t)c * np + r];
            R xr = xr_[c], xi = xi_[c];
            ar = fma(qr, xr, ar); ar = fma(-qi, xi, ar);
            ai = fma(qr, xi, ai); ai = fma(qi, xr, ai);
          }
          V[(size_t)k * np + r] = ar;
          V[vb + (size_t)k * np + r] = ai;
        }
        __syncthreads();
      }
      // 5c. f(kB+j) = u_j . v_k  (accumulated in FP64)
      const int64_t npts = p.nsteps + 1;
      double* out = p.per_or + 2 * (size_t)o * npts;
      for (int64_t t = tid; t < npts; t += kThreads) {
        int j = (int)(t % p.block), k = (int)(t / p.block);
        const R* ur = U + (size_t)j * np;
        const R* ui = U + ub + (size_t)j * np;
        const R* vr = V + (size_t)k * np;
        const R* vi = V + vb + (size_t)k * np;
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < np; r++) {
          double a = (double)ur[r], b = (double)ui[r], c = (double)vr[r], d = (double)vi[r];
          sr += a * c - b * d;
          si += a * d + b * c;
        }
        out[2 * t] = sr;
        out[2 * t + 1] = si;
      }
      __syncthreads();
    }
  }
}

// Grid-weighted orientation sum in grid order, FP64 (powder.m:280-286).
__global__ void k_weighted_sum(const double* __restrict__ per_or, const double* __restrict__ weights,
                               int64_t or_off, int64_t n_or, int64_t npts, double* __restrict__ fid) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const double* base = per_or + 2 * (size_t)or_off * npts;
  double sr = weights[or_off] * base[2 * t];
  double si = weights[or_off] * base[2 * t + 1];
  for (int64_t k = 1; k < n_or; k++) {
    double w = weights[or_off + k];
    sr += w * base[2 * ((size_t)k * npts + t)];
    si += w * base[2 * ((size_t)k * npts + t) + 1];
  }
  fid[2 * t] = sr;
  fid[2 * t + 1] = si;
}

}  // namespace spk

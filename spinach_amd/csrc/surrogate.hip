// Learned powder interpolation: a Fourier-feature MLP f(alpha, beta, gamma, t)
// -> (Re fid, Im fid), trained on FIDs of the plain method at a coarse grid
// and evaluated on a fine grid with the powder sum formed on the device.
//
// No counterpart in the reference (Spinach has no learned interpolation); the
// plain method stays the acceptance check and the fallback (see surrogate.py).
//
// Network: F input features -> depth hidden layers of width W with tanh ->
// 2 linear outputs. All matrix products are BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; master weights, Adam
// moments and the loss are FP32; the powder sum is FP64.
//
// Features of a point (orientation o, time index j, t = j*dt, tau = t/T with
// T the record length, s = t/Tf with Tf the time scale of the harmonics):
//   cos(k a), sin(k a), cos(k b), sin(k b), cos(k g), sin(k g),  k = 1..KA
//   (3 cos^2 b - 1)/2, sin b cos b cos g, sin b cos b sin g, sin^2 b cos 2g, sin^2 b sin 2g
//       (up to constant factors the real and imaginary parts of the Wigner
//       row D2(0,m)(a,b,g), on which a secular rank-2 Hamiltonian depends)
//   cos(pi m s), sin(pi m s),                                      m = 1..KT
//   tau, 1
// With Tf = KT*dt the harmonics reach the Nyquist frequency 1/(2 dt) in steps
// of 1/(2 KT dt), spanning the band of the FID; with Tf = T they span only
// KT/(2T).
// zero-padded to F (a multiple of 32).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spinach_powder.h"

namespace sgk {

typedef unsigned short bf16;
typedef short s8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16 f2bf(float f) {
  // round to nearest even
  unsigned u;
  memcpy(&u, &f, 4);
  unsigned r = ((u >> 16) & 1u) + 0x7fffu;
  return (bf16)((u + r) >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  unsigned u = ((unsigned)h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// ------------------------------------------------------------------ features
struct FeatArgs {
  const double* alphas;
  const double* betas;
  const double* gammas;
  const int32_t* pt_or;   // orientation of each point (nullptr: o = p / npts)
  const int32_t* pt_t;    // time index of each point (nullptr: j = p % npts)
  int64_t p0;             // first point of this chunk
  int64_t n_points;
  int64_t npts;
  double dt;
  double T;
  double Tf;
  int ka, kt, F;
  bf16* X;  // n_points x F
};

__global__ void k_features(FeatArgs a) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_points) return;
  int64_t g = a.p0 + p;
  int64_t o = a.pt_or ? a.pt_or[g] : g / a.npts;
  int64_t j = a.pt_t ? a.pt_t[g] : g % a.npts;
  double ang[3] = {a.alphas[o], a.betas[o], a.gammas[o]};
  double tau = (double)j * a.dt / a.T;
  double sf = (double)j * a.dt / a.Tf;
  bf16* x = a.X + p * a.F;
  int c = 0;
  for (int d = 0; d < 3; d++)
    for (int k = 1; k <= a.ka; k++) {
      x[c++] = f2bf((float)cos(k * ang[d]));
      x[c++] = f2bf((float)sin(k * ang[d]));
    }
  {
    const double cb = cos(ang[1]), sb = sin(ang[1]);
    x[c++] = f2bf((float)(0.5 * (3.0 * cb * cb - 1.0)));
    x[c++] = f2bf((float)(sb * cb * cos(ang[2])));
    x[c++] = f2bf((float)(sb * cb * sin(ang[2])));
    x[c++] = f2bf((float)(sb * sb * cos(2.0 * ang[2])));
    x[c++] = f2bf((float)(sb * sb * sin(2.0 * ang[2])));
  }
  for (int m = 1; m <= a.kt; m++) {
    double ph = 3.14159265358979323846 * m * sf;
    x[c++] = f2bf((float)cos(ph));
    x[c++] = f2bf((float)sin(ph));
  }
  x[c++] = f2bf((float)tau);
  x[c++] = f2bf(1.0f);
  for (; c < a.F; c++) x[c] = 0;
}

// ------------------------------------------------------------- BF16 GEMM
// C(MxN) = op(A)(MxK) * op(B)(KxN) in FP32, with A, B in BF16.
//   TA = false: A stored M x K (row-major, lda); true: stored K x M.
//   TB = false: B stored K x N (row-major, ldb); true: stored N x K.
// Epilogue EPI: 0 = store C (fp32, ldc); 1 = +bias[n], tanh, store fp32 and a
// BF16 copy; 2 = +bias[n], store fp32; 3 = accumulate into C (C += A*B);
// 4 = +bias[n], tanh, store the BF16 copy only (inference: no backward pass
// reads the FP32 activations).
// Tile 64x64 per workgroup (4 waves, 32x32 each as 2x2 MFMA tiles), k-panels
// of 32. M, N and K arbitrary (masked). Split K (EPI 0 only): workgroup z of
// gridDim.z sums k in [z*kchunk, (z+1)*kchunk) and stores its partial product
// at C + z*zstride; k_sum_parts adds the partials in z order.
struct GemmArgs {
  const bf16* A;
  const bf16* B;
  float* C;
  bf16* Cb;
  const float* bias;
  int M, N, K, lda, ldb, ldc;
  int kchunk;       // k range of one split (a multiple of 32), K when not split
  int64_t zstride;  // floats between the partial products of two splits
};

// stage one 64 x 32 operand panel into LDS: rows of the panel are output rows
// (A: i, B: j), columns are k. CONTIG: the 32 k-values of a row are contiguous
// in memory (A row-major M x K, or B stored N x K), read 8 at a time (16 bytes)
// when the row stride and the panel start allow it.
template <bool CONTIG>
__device__ inline void stage_panel(bf16 (*S)[40], const bf16* X, int ld, int r0, int rmax, int k0, int kend,
                                   bool vec) {
  constexpr int KP = 32;
  const int tid = threadIdx.x;
  if (CONTIG && vec) {
    // 64 rows x 4 groups of 8: one 16-byte load per thread
    const int r = tid >> 2, c = 8 * (tid & 3);
    const int gr = r0 + r;
    s8v v = s8v{0, 0, 0, 0, 0, 0, 0, 0};
    if (gr < rmax && k0 + c + 8 <= kend) {
      v = *reinterpret_cast<const s8v*>(X + (size_t)gr * ld + k0 + c);
    } else if (gr < rmax) {
      for (int u = 0; u < 8; u++)
        if (k0 + c + u < kend) v[u] = (short)X[(size_t)gr * ld + k0 + c + u];
    }
    *reinterpret_cast<s8v*>(&S[r][c]) = v;
    return;
  }
  for (int e = tid; e < 64 * KP; e += 256) {
    int r, c;
    if (CONTIG) { r = e / KP; c = e % KP; } else { c = e / 64; r = e % 64; }
    const int gr = r0 + r;
    bf16 v = 0;
    if (gr < rmax && k0 + c < kend) v = CONTIG ? X[(size_t)gr * ld + k0 + c] : X[(size_t)(k0 + c) * ld + gr];
    S[r][c] = v;
  }
}

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  constexpr int KP = 32, LDK = KP + 8;  // padded rows (16-byte aligned, bank spread)
  __shared__ __attribute__((aligned(16))) bf16 As[64][LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[64][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
  float* C = g.C + (size_t)blockIdx.z * g.zstride;
  // 16-byte staging needs 8-element aligned rows
  const bool va = !TA && (g.lda % 8) == 0, vb = TB && (g.ldb % 8) == 0;
  f4v acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f4v{0, 0, 0, 0};
  for (int k0 = kbeg; k0 < kend; k0 += KP) {
    stage_panel<!TA>(As, g.A, g.lda, i0, g.M, k0, kend, va);
    stage_panel<TB>(Bs, g.B, g.ldb, j0, g.N, k0, kend, vb);
    __syncthreads();
    s8v xa[2], xb[2];
    for (int t = 0; t < 2; t++) {
      // 16-byte aligned: rows are 80 bytes, the column offset is 16 bytes
      xa[t] = *reinterpret_cast<const s8v*>(&As[32 * wr + 16 * t + (lane & 15)][8 * (lane >> 4)]);
      xb[t] = *reinterpret_cast<const s8v*>(&Bs[32 * wc + 16 * t + (lane & 15)][8 * (lane >> 4)]);
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[a], xb[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 4; e++) {
        int i = i0 + 32 * wr + 16 * a + 4 * (lane >> 4) + e;
        int j = j0 + 32 * wc + 16 * b + (lane & 15);
        if (i >= g.M || j >= g.N) continue;
        float v = acc[a][b][e];
        size_t o = (size_t)i * g.ldc + j;
        if (EPI == 0) {
          C[o] = v;
        } else if (EPI == 1) {
          v = tanhf(v + g.bias[j]);
          C[o] = v;
          g.Cb[o] = f2bf(v);
        } else if (EPI == 2) {
          C[o] = v + g.bias[j];
        } else if (EPI == 4) {
          g.Cb[o] = f2bf(tanhf(v + g.bias[j]));
        } else {
          C[o] += v;
        }
      }
}

// ------------------------------------------------------------ elementwise
// output layer gradient: dY = 2 (Y - target) / n, loss partial sums per block
__global__ void k_loss_grad(const float* Y, const float* target, int64_t n, float* dY, bf16* dYb, float* loss_part) {
  __shared__ float red[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < 2 * n) {
    float d = Y[i] - target[i];
    l = d * d;
    float gr = 2.f * d / (float)n;
    dY[i] = gr;
    dYb[i] = f2bf(gr);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (unsigned s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ = dH * (1 - H^2) for a tanh layer; BF16 copy for the next products
__global__ void k_tanh_bwd(const float* dH, const float* H, int64_t n, float* dZ, bf16* dZb) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float h = H[i];
  float v = dH[i] * (1.f - h * h);
  dZ[i] = v;
  dZb[i] = f2bf(v);
}

// column sums (bias gradient), stage 1: workgroup (x, y) sums rows
// y, y + gridDim.y, ... of columns [64x, 64x + 64) into part[y][j]; the 256
// threads are 64 columns x 4 row phases, combined in LDS in a fixed order
__global__ void __launch_bounds__(256) k_colsum_part(const float* dZ, int64_t rows, int cols, float* part) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  float s = 0.f;
  if (j < cols)
    for (int64_t i = (int64_t)blockIdx.y * 4 + ph; i < rows; i += (int64_t)gridDim.y * 4) s += dZ[i * cols + j];
  red[ph][c] = s;
  __syncthreads();
  if (ph == 0 && j < cols) part[(size_t)blockIdx.y * cols + j] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

// out[i] = sum_z part[z * n + i], z in order (split-K products, column-sum partials)
__global__ void k_sum_parts(const float* part, int nz, int64_t n, float* out) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nz; z++) s += part[(size_t)z * n + i];
  out[i] = s;
}

// Adam on FP32 master weights, BF16 copy refreshed
__global__ void k_adam(float* w, const float* gr, float* m, float* v, bf16* wb, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float g = gr[i];
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  float wi = w[i] - upd;
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// powder accumulation of a chunk of whole orientations: fid[j] += sum_o w_o Y[o][j]
// (orientation order inside the chunk, FP64)
__global__ void k_powder_accum(const float* Y, const double* weights, int64_t o0, int64_t n_or, int64_t npts,
                               double* fid) {
  int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= npts) return;
  double sr = 0, si = 0;
  for (int64_t o = 0; o < n_or; o++) {
    const float* y = Y + 2 * ((size_t)o * npts + j);
    double w = weights[o0 + o];
    sr += w * (double)y[0];
    si += w * (double)y[1];
  }
  fid[2 * j] += sr;
  fid[2 * j + 1] += si;
}

}  // namespace sgk

// ----------------------------------------------------------------- host side
namespace {

struct SgErr : std::runtime_error {
  int code;
  SgErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SGCHK(x)                                                                                  \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw SgErr(SP_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Dev {
  void* p = nullptr;
  size_t n = 0;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p && n >= b) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(b, 16));
    if (e != hipSuccess) {
      p = nullptr;
      throw SgErr(SP_ENOMEM, "hipMalloc failed in the surrogate");
    }
    n = std::max<size_t>(b, 16);
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace

struct sp_surrogate {
  int device = 0;
  hipStream_t stream = nullptr;
  int ka = 3, kt = 22, F = 64, W = 256, depth = 4;
  double T = 1.0, Tf = 1.0;
  std::string err;
  int code = 0;
  // parameters: layer l maps fan_in(l) -> fan_out(l); W_l stored out x in (row-major)
  std::vector<int> fin, fout;
  std::vector<size_t> woff, boff;
  size_t nparam = 0;
  Dev w32, wbf, m, v, g;  // FP32 master, BF16 copy, Adam moments, gradient
  int64_t step = 0;
  // activations for a batch of up to `cap` points
  int64_t cap = 0;
  Dev X, Hf, Hb, dHf, dHb, Y, dY, dYb, tgt, lossp, or_idx, t_idx, grid;
  Dev part;  // split-K partial products and column-sum partials of the backward pass
};

namespace {

int sg_fail(sp_surrogate* s, int code, const std::string& m) {
  if (s) {
    s->err = m;
    s->code = code;
  }
  return code;
}

constexpr int kMaxSplit = 32;  // split-K factor and column-sum partial rows of the backward pass

void ensure_batch(sp_surrogate* s, int64_t n) {
  if (n <= s->cap) return;
  s->cap = n;
  size_t wmax = (size_t)std::max(s->W, s->F);
  s->X.alloc((size_t)n * s->F * sizeof(sgk::bf16));
  s->Hf.alloc((size_t)s->depth * n * s->W * sizeof(float));
  s->Hb.alloc((size_t)s->depth * n * s->W * sizeof(sgk::bf16));
  s->dHf.alloc((size_t)2 * n * wmax * sizeof(float));
  s->dHb.alloc((size_t)2 * n * wmax * sizeof(sgk::bf16));
  s->Y.alloc((size_t)n * 2 * sizeof(float));
  s->dY.alloc((size_t)n * 2 * sizeof(float));
  s->dYb.alloc((size_t)n * 2 * sizeof(sgk::bf16));
  s->tgt.alloc((size_t)n * 2 * sizeof(float));
  s->lossp.alloc((size_t)((2 * n + 255) / 256) * sizeof(float));
  s->part.alloc((size_t)kMaxSplit * wmax * wmax * sizeof(float));
}

template <bool TA, bool TB, int EPI>
void gemm(sp_surrogate* s, const sgk::bf16* A, int lda, const sgk::bf16* B, int ldb, float* C, int ldc,
          sgk::bf16* Cb, const float* bias, int M, int N, int K) {
  sgk::GemmArgs a{A, B, C, Cb, bias, M, N, K, lda, ldb, ldc, K, 0};
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL((sgk::k_gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, s->stream, a);
  SGCHK(hipGetLastError());
}

// C (M x N, ldc = N) = op(A) op(B) with the K dimension split so that a few
// output tiles still fill the GPU (the weight gradients: M, N <= 256, K = the
// batch); partials in s->part, summed in a fixed order
template <bool TA, bool TB>
void gemm_splitk(sp_surrogate* s, const sgk::bf16* A, int lda, const sgk::bf16* B, int ldb, float* C, int M, int N,
                 int K) {
  const int tiles = ((N + 63) / 64) * ((M + 63) / 64);
  int nz = std::max(1, std::min(kMaxSplit, std::min(K / 512, 256 / std::max(1, tiles))));
  const int kchunk = ((K + nz - 1) / nz + 31) / 32 * 32;
  nz = (K + kchunk - 1) / kchunk;
  if (nz <= 1) {
    gemm<TA, TB, 0>(s, A, lda, B, ldb, C, N, nullptr, nullptr, M, N, K);
    return;
  }
  const int64_t mn = (int64_t)M * N;
  sgk::GemmArgs a{A, B, s->part.as<float>(), nullptr, nullptr, M, N, K, lda, ldb, N, kchunk, mn};
  dim3 grid((N + 63) / 64, (M + 63) / 64, nz);
  hipLaunchKernelGGL((sgk::k_gemm_bf16<TA, TB, 0>), grid, dim3(256), 0, s->stream, a);
  SGCHK(hipGetLastError());
  hipLaunchKernelGGL(sgk::k_sum_parts, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s->stream,
                     s->part.as<float>(), nz, mn, C);
  SGCHK(hipGetLastError());
}

// db[j] = sum_i dZ[i][j] over n rows of `cols` columns
void colsum(sp_surrogate* s, const float* dz, int64_t n, int cols, float* db) {
  const int ny = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxSplit, (n + 255) / 256));
  hipLaunchKernelGGL(sgk::k_colsum_part, dim3((cols + 63) / 64, ny), dim3(256), 0, s->stream, dz, n, cols,
                     s->part.as<float>());
  SGCHK(hipGetLastError());
  hipLaunchKernelGGL(sgk::k_sum_parts, dim3((cols + 255) / 256), dim3(256), 0, s->stream, s->part.as<float>(), ny,
                     (int64_t)cols, db);
  SGCHK(hipGetLastError());
}

// forward pass of n points whose features are in s->X; output in s->Y (n x 2);
// train: keep the FP32 activations for backward()
void forward(sp_surrogate* s, int64_t n, bool train = true) {
  const sgk::bf16* in = s->X.as<sgk::bf16>();
  int ld_in = s->F;
  float* Hf = s->Hf.as<float>();
  sgk::bf16* Hb = s->Hb.as<sgk::bf16>();
  const float* w32 = s->w32.as<float>();
  const sgk::bf16* wb = s->wbf.as<sgk::bf16>();
  for (int l = 0; l < s->depth; l++) {
    float* h = Hf + (size_t)l * s->cap * s->W;
    sgk::bf16* hb = Hb + (size_t)l * s->cap * s->W;
    if (train)
      gemm<false, true, 1>(s, in, ld_in, wb + s->woff[l], s->fin[l], h, s->W, hb, w32 + s->boff[l], (int)n, s->W,
                           s->fin[l]);
    else
      gemm<false, true, 4>(s, in, ld_in, wb + s->woff[l], s->fin[l], nullptr, s->W, hb, w32 + s->boff[l], (int)n,
                           s->W, s->fin[l]);
    in = hb;
    ld_in = s->W;
  }
  int L = s->depth;
  gemm<false, true, 2>(s, in, ld_in, wb + s->woff[L], s->fin[L], s->Y.as<float>(), 2, nullptr, w32 + s->boff[L],
                       (int)n, 2, s->fin[L]);
}

// backward pass after forward(); targets in s->tgt; gradient into s->g
float backward(sp_surrogate* s, int64_t n) {
  const int L = s->depth;
  int nb = (int)((2 * n + 255) / 256);
  hipLaunchKernelGGL(sgk::k_loss_grad, dim3(nb), dim3(256), 0, s->stream, s->Y.as<float>(), s->tgt.as<float>(), n,
                     s->dY.as<float>(), s->dYb.as<sgk::bf16>(), s->lossp.as<float>());
  SGCHK(hipGetLastError());
  float* gw = s->g.as<float>();
  const sgk::bf16* wb = s->wbf.as<sgk::bf16>();
  float* Hf = s->Hf.as<float>();
  sgk::bf16* Hb = s->Hb.as<sgk::bf16>();
  // dZ of the current layer (fp32 + bf16), starting with the output layer
  const float* dz = s->dY.as<float>();
  const sgk::bf16* dzb = s->dYb.as<sgk::bf16>();
  int nout = 2;
  size_t wmax = (size_t)std::max(s->W, s->F);
  for (int l = L; l >= 0; l--) {
    const sgk::bf16* inb = (l == 0) ? s->X.as<sgk::bf16>() : Hb + (size_t)(l - 1) * s->cap * s->W;
    const int fin = s->fin[l];
    // dW_l = dZ^T (nout x n) * In (n x fin)
    gemm_splitk<true, false>(s, dzb, nout, inb, fin, gw + s->woff[l], nout, fin, (int)n);
    colsum(s, dz, n, nout, gw + s->boff[l]);
    if (l == 0) break;
    // dH_{l-1} = dZ (n x nout) * W_l (nout x fin)
    float* dh = s->dHf.as<float>() + ((l & 1) ? 0 : (size_t)n * wmax);
    sgk::bf16* dhb = s->dHb.as<sgk::bf16>() + ((l & 1) ? 0 : (size_t)n * wmax);
    gemm<false, false, 0>(s, dzb, nout, wb + s->woff[l], fin, dh, fin, nullptr, nullptr, (int)n, fin, nout);
    const float* H = Hf + (size_t)(l - 1) * s->cap * s->W;
    int64_t cnt = n * fin;
    hipLaunchKernelGGL(sgk::k_tanh_bwd, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s->stream, dh, H, cnt, dh,
                       dhb);
    SGCHK(hipGetLastError());
    dz = dh;
    dzb = dhb;
    nout = fin;
  }
  std::vector<float> lp(nb);
  SGCHK(hipMemcpyAsync(lp.data(), s->lossp.p, nb * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  SGCHK(hipStreamSynchronize(s->stream));
  double tot = 0;
  for (float x : lp) tot += x;
  return (float)(tot / (double)n);
}

void features(sp_surrogate* s, const double* a, const double* b, const double* gm, const int32_t* por,
              const int32_t* pt, int64_t p0, int64_t n, int64_t npts, double dt) {
  sgk::FeatArgs fa{a, b, gm, por, pt, p0, n, npts, dt, s->T, s->Tf, s->ka, s->kt, s->F, s->X.as<sgk::bf16>()};
  hipLaunchKernelGGL(sgk::k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, fa);
  SGCHK(hipGetLastError());
}

}  // namespace

extern "C" {

sp_surrogate* sp_surrogate_create(int device, int width, int depth, int ka, int kt, double t_max, double t_feat,
                                  uint64_t seed) {
  std::unique_ptr<sp_surrogate> s(new sp_surrogate);
  try {
    if (width < 32 || width % 32 || depth < 1 || depth > 16 || ka < 0 || kt < 0 || !(t_max > 0) || !(t_feat > 0))
      throw SgErr(SP_EINVAL, "invalid surrogate shape");
    SGCHK(hipSetDevice(device));
    s->device = device;
    SGCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->W = width;
    s->depth = depth;
    s->ka = ka;
    s->kt = kt;
    s->T = t_max;
    s->Tf = t_feat;
    int nf = 6 * ka + 5 + 2 * kt + 2;
    s->F = (nf + 31) / 32 * 32;
    size_t off = 0;
    for (int l = 0; l <= depth; l++) {
      int fi = (l == 0) ? s->F : width, fo = (l == depth) ? 2 : width;
      s->fin.push_back(fi);
      s->fout.push_back(fo);
      s->woff.push_back(off);
      off += (size_t)fi * fo;
      s->boff.push_back(off);
      off += fo;
    }
    s->nparam = off;
    std::vector<float> w(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l <= depth; l++) {
      std::normal_distribution<float> nd(0.f, std::sqrt(1.0f / (float)s->fin[l]));
      for (size_t i = 0; i < (size_t)s->fin[l] * s->fout[l]; i++) w[s->woff[l] + i] = nd(rng);
    }
    std::vector<sgk::bf16> wb(off);
    for (size_t i = 0; i < off; i++) wb[i] = sgk::f2bf(w[i]);
    s->w32.alloc(off * 4);
    s->wbf.alloc(off * 2);
    s->m.alloc(off * 4);
    s->v.alloc(off * 4);
    s->g.alloc(off * 4);
    SGCHK(hipMemcpy(s->w32.p, w.data(), off * 4, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(s->wbf.p, wb.data(), off * 2, hipMemcpyHostToDevice));
    SGCHK(hipMemset(s->m.p, 0, off * 4));
    SGCHK(hipMemset(s->v.p, 0, off * 4));
    return s.release();
  } catch (std::exception& e) {
    return nullptr;
  }
}

void sp_surrogate_destroy(sp_surrogate* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* sp_surrogate_last_error(const sp_surrogate* s) { return s ? s->err.c_str() : "no surrogate"; }

int64_t sp_surrogate_num_params(const sp_surrogate* s) { return s ? (int64_t)s->nparam : 0; }

int sp_surrogate_get_params(sp_surrogate* s, float* out) {
  if (!s || !out) return SP_EINVAL;
  try {
    SGCHK(hipSetDevice(s->device));
    SGCHK(hipMemcpy(out, s->w32.p, s->nparam * 4, hipMemcpyDeviceToHost));
    return SP_OK;
  } catch (SgErr& e) {
    return sg_fail(s, e.code, e.what());
  }
}

int sp_surrogate_set_params(sp_surrogate* s, const float* in) {
  if (!s || !in) return SP_EINVAL;
  try {
    SGCHK(hipSetDevice(s->device));
    std::vector<sgk::bf16> wb(s->nparam);
    for (size_t i = 0; i < s->nparam; i++) wb[i] = sgk::f2bf(in[i]);
    SGCHK(hipMemcpy(s->w32.p, in, s->nparam * 4, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(s->wbf.p, wb.data(), s->nparam * 2, hipMemcpyHostToDevice));
    return SP_OK;
  } catch (SgErr& e) {
    return sg_fail(s, e.code, e.what());
  }
}

// Loss and gradient for explicit points (orientation angles per point, time
// index per point, targets n x 2 interleaved): used to check the kernels.
int sp_surrogate_grad(sp_surrogate* s, const double* a, const double* b, const double* gm, const int32_t* tidx,
                      int64_t n, double dt, const float* target, float* loss, float* grad) {
  if (!s || !a || !b || !gm || !tidx || !target || n <= 0) return SP_EINVAL;
  try {
    SGCHK(hipSetDevice(s->device));
    ensure_batch(s, n);
    Dev da, dbv, dg, dt_, dor;
    da.alloc(n * 8);
    dbv.alloc(n * 8);
    dg.alloc(n * 8);
    dt_.alloc(n * 4);
    dor.alloc(n * 4);
    std::vector<int32_t> oi(n);
    for (int64_t i = 0; i < n; i++) oi[i] = (int32_t)i;
    SGCHK(hipMemcpy(da.p, a, n * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dbv.p, b, n * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dg.p, gm, n * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dt_.p, tidx, n * 4, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dor.p, oi.data(), n * 4, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(s->tgt.p, target, n * 8, hipMemcpyHostToDevice));
    features(s, da.as<double>(), dbv.as<double>(), dg.as<double>(), dor.as<int32_t>(), dt_.as<int32_t>(), 0, n, 1, dt);
    forward(s, n);
    float l = backward(s, n);
    if (loss) *loss = l;
    if (grad) SGCHK(hipMemcpy(grad, s->g.p, s->nparam * 4, hipMemcpyDeviceToHost));
    return SP_OK;
  } catch (SgErr& e) {
    return sg_fail(s, e.code, e.what());
  }
}

// Train on FIDs fids[o][j] (n_or x npts interleaved complex) at the given
// orientations: `steps` Adam steps on random mini-batches of `batch` points.
int sp_surrogate_train(sp_surrogate* s, const double* a, const double* b, const double* gm, int64_t n_or,
                       const double* fids, int64_t npts, double dt, int64_t steps, int64_t batch, double lr,
                       uint64_t seed, float* loss_trace) {
  if (!s || !a || !b || !gm || !fids || n_or <= 0 || npts <= 0 || steps < 0 || batch <= 0) return SP_EINVAL;
  try {
    SGCHK(hipSetDevice(s->device));
    ensure_batch(s, batch);
    Dev da, dbv, dg, dor, dt_;
    da.alloc(n_or * 8);
    dbv.alloc(n_or * 8);
    dg.alloc(n_or * 8);
    SGCHK(hipMemcpy(da.p, a, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dbv.p, b, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dg.p, gm, n_or * 8, hipMemcpyHostToDevice));
    dor.alloc(batch * 4);
    dt_.alloc(batch * 4);
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int64_t> uo(0, n_or - 1), ut(0, npts - 1);
    std::vector<int32_t> oi(batch), ti(batch);
    std::vector<float> tg(2 * batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t nblk = (int64_t)((s->nparam + 255) / 256);
    for (int64_t it = 0; it < steps; it++) {
      for (int64_t i = 0; i < batch; i++) {
        int64_t o = uo(rng), j = ut(rng);
        oi[i] = (int32_t)o;
        ti[i] = (int32_t)j;
        tg[2 * i] = (float)fids[2 * (o * npts + j)];
        tg[2 * i + 1] = (float)fids[2 * (o * npts + j) + 1];
      }
      SGCHK(hipMemcpyAsync(dor.p, oi.data(), batch * 4, hipMemcpyHostToDevice, s->stream));
      SGCHK(hipMemcpyAsync(dt_.p, ti.data(), batch * 4, hipMemcpyHostToDevice, s->stream));
      SGCHK(hipMemcpyAsync(s->tgt.p, tg.data(), batch * 8, hipMemcpyHostToDevice, s->stream));
      features(s, da.as<double>(), dbv.as<double>(), dg.as<double>(), dor.as<int32_t>(), dt_.as<int32_t>(), 0, batch,
               npts, dt);
      forward(s, batch);
      float l = backward(s, batch);
      if (loss_trace) loss_trace[it] = l;
      s->step++;
      float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
      hipLaunchKernelGGL(sgk::k_adam, dim3((unsigned)nblk), dim3(256), 0, s->stream, s->w32.as<float>(),
                         s->g.as<float>(), s->m.as<float>(), s->v.as<float>(), s->wbf.as<sgk::bf16>(),
                         (int64_t)s->nparam, (float)lr, b1, b2, eps, c1, c2);
      SGCHK(hipGetLastError());
    }
    SGCHK(hipStreamSynchronize(s->stream));
    return SP_OK;
  } catch (SgErr& e) {
    return sg_fail(s, e.code, e.what());
  }
}

// Evaluate the network on a grid: per-orientation FIDs (optional, n_or x npts
// interleaved complex) and the weighted powder sum (fid_out, FP64). kernel_ms
// receives the device time of the evaluation (HIP events).
int sp_surrogate_powder(sp_surrogate* s, const double* a, const double* b, const double* gm, const double* w,
                        int64_t n_or, int64_t npts, double dt, double* fid_out, double* per_orient_out,
                        double* kernel_ms) {
  if (!s || !a || !b || !gm || !w || n_or <= 0 || npts <= 0 || !fid_out) return SP_EINVAL;
  try {
    SGCHK(hipSetDevice(s->device));
    const int64_t chunk_or = std::max<int64_t>(1, (1 << 20) / npts);  // ~1M points per chunk
    ensure_batch(s, chunk_or * npts);
    Dev da, dbv, dg, dw, dfid;
    da.alloc(n_or * 8);
    dbv.alloc(n_or * 8);
    dg.alloc(n_or * 8);
    dw.alloc(n_or * 8);
    dfid.alloc(npts * 16);
    SGCHK(hipMemcpy(da.p, a, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dbv.p, b, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dg.p, gm, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemcpy(dw.p, w, n_or * 8, hipMemcpyHostToDevice));
    SGCHK(hipMemset(dfid.p, 0, npts * 16));
    hipEvent_t e0, e1;
    SGCHK(hipEventCreate(&e0));
    SGCHK(hipEventCreate(&e1));
    SGCHK(hipEventRecord(e0, s->stream));
    for (int64_t o0 = 0; o0 < n_or; o0 += chunk_or) {
      int64_t no = std::min(chunk_or, n_or - o0);
      int64_t n = no * npts;
      features(s, da.as<double>(), dbv.as<double>(), dg.as<double>(), nullptr, nullptr, o0 * npts, n, npts, dt);
      forward(s, n, false);
      hipLaunchKernelGGL(sgk::k_powder_accum, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s->stream,
                         s->Y.as<float>(), dw.as<double>(), o0, no, npts, dfid.as<double>());
      SGCHK(hipGetLastError());
      if (per_orient_out) {
        std::vector<float> y(2 * n);
        SGCHK(hipMemcpyAsync(y.data(), s->Y.p, 2 * n * 4, hipMemcpyDeviceToHost, s->stream));
        SGCHK(hipStreamSynchronize(s->stream));
        for (int64_t i = 0; i < 2 * n; i++) per_orient_out[2 * o0 * npts + i] = y[i];
      }
    }
    SGCHK(hipEventRecord(e1, s->stream));
    SGCHK(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    SGCHK(hipMemcpy(fid_out, dfid.p, npts * 16, hipMemcpyDeviceToHost));
    return SP_OK;
  } catch (SgErr& e) {
    return sg_fail(s, e.code, e.what());
  }
}

}  // extern "C"

// refcpu: CPU FP64 restatement of Spinach's powder-averaged time-domain
// propagation path. TEST INFRASTRUCTURE ONLY: this is the oracle the HIP path
// is checked against and the CPU baseline that bench.py times beside it. The
// product library never links or calls it.
//
// Reference algorithm (paths relative to the Spinach root):
//   kernel/contexts/powder.m:226-286   orientation loop, H=I+orientation(Q,Omega),
//                                      H=(H+H')/2, weighted sum in grid order
//   kernel/orientation.m:29-59         H = sum D^r(k,m) Q{r}{k,m}, then (H+H')/2
//   kernel/utilities/wigner.m:39-77    rank-1/rank-2 Wigner matrices, descending m
//   experiments/acquire.m:42-73        L=H+iR+iK (R=K=0), evolution(...,'observable')
//   kernel/evolution.m:121-163,503-587 reduce(L',coil); per subspace L_sub=P'LP,
//                                      clean_up, propagator, detection loop
//   kernel/reduce.m:127-223            ZTE then path tracing
//   kernel/utilities/zte.m:43-142      zero track elimination
//   kernel/utilities/path_trace.m:36-183, scomponents.m, binpack.m
//   kernel/propagator.m:34-223         expm below small_matrix, else chop-Taylor
//                                      with squaring
//   kernel/utilities/clean_up.m:22-90  tol*round(A/tol)
//   kernel/utilities/cheap_norm.m:34-50  1-norm on CPU
//   kernel/step.m:204-375              reordered Taylor stepping
//   kernel/krylov.m:139-163            step()-based trajectory for dim>=krylov_tol
//
// MATLAB's expm is closed; the dense exponential here is the published
// Higham (2005) scaling-and-squaring Pade algorithm ("The scaling and
// squaring method for the matrix exponential revisited", SIAM J. Matrix
// Anal. Appl. 26:1179), which is what MATLAB documents for expm.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using cd = std::complex<double>;
using i64 = int64_t;

namespace {

thread_local std::string g_err;

struct Csr {
  i64 n = 0;
  std::vector<i64> ptr, idx;
  std::vector<cd> val;
};

// --------------------------------------------------------------- dense helpers
struct ZMat {  // dense n x n, row-major, split real/imaginary planes
  int n = 0;
  std::vector<double> re, im;
  ZMat() = default;
  explicit ZMat(int n_) : n(n_), re((size_t)n_ * n_, 0.0), im((size_t)n_ * n_, 0.0) {}
  static ZMat eye(int n_) {
    ZMat m(n_);
    for (int i = 0; i < n_; i++) m.re[(size_t)i * n_ + i] = 1.0;
    return m;
  }
};

// C = A*B
static ZMat matmul(const ZMat& a, const ZMat& b) {
  int n = a.n;
  ZMat c(n);
  for (int i = 0; i < n; i++) {
    double* cr = &c.re[(size_t)i * n];
    double* ci = &c.im[(size_t)i * n];
    for (int k = 0; k < n; k++) {
      double ar = a.re[(size_t)i * n + k], ai = a.im[(size_t)i * n + k];
      if (ar == 0.0 && ai == 0.0) continue;
      const double* br = &b.re[(size_t)k * n];
      const double* bi = &b.im[(size_t)k * n];
      for (int j = 0; j < n; j++) {
        cr[j] += ar * br[j] - ai * bi[j];
        ci[j] += ar * bi[j] + ai * br[j];
      }
    }
  }
  return c;
}

static double norm1(const ZMat& a) {  // max column sum of |a_ij|
  int n = a.n;
  std::vector<double> col(n, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) col[j] += std::hypot(a.re[(size_t)i * n + j], a.im[(size_t)i * n + j]);
  double m = 0;
  for (double v : col) m = std::max(m, v);
  return m;
}

static void axpy(ZMat& y, double s, const ZMat& x) {  // y += s*x (real s)
  for (size_t i = 0; i < y.re.size(); i++) {
    y.re[i] += s * x.re[i];
    y.im[i] += s * x.im[i];
  }
}

static void scale(ZMat& y, double s) {
  for (size_t i = 0; i < y.re.size(); i++) {
    y.re[i] *= s;
    y.im[i] *= s;
  }
}

// Solve A X = B in place (B overwritten by X); Gaussian elimination with
// partial pivoting on complex matrices.
static void solve(ZMat a, ZMat& b) {
  int n = a.n;
  auto A = [&](int i, int j) -> cd { return cd(a.re[(size_t)i * n + j], a.im[(size_t)i * n + j]); };
  auto setA = [&](int i, int j, cd v) {
    a.re[(size_t)i * n + j] = v.real();
    a.im[(size_t)i * n + j] = v.imag();
  };
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::abs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (std::abs(A(i, k)) > best) {
        best = std::abs(A(i, k));
        p = i;
      }
    if (best == 0.0) throw std::runtime_error("singular matrix in Pade solve");
    if (p != k) {
      for (int j = 0; j < n; j++) {
        std::swap(a.re[(size_t)k * n + j], a.re[(size_t)p * n + j]);
        std::swap(a.im[(size_t)k * n + j], a.im[(size_t)p * n + j]);
        std::swap(b.re[(size_t)k * n + j], b.re[(size_t)p * n + j]);
        std::swap(b.im[(size_t)k * n + j], b.im[(size_t)p * n + j]);
      }
    }
    cd piv = A(k, k);
    for (int i = k + 1; i < n; i++) {
      cd f = A(i, k) / piv;
      if (f == cd(0)) continue;
      setA(i, k, cd(0));
      for (int j = k + 1; j < n; j++) setA(i, j, A(i, j) - f * A(k, j));
      double fr = f.real(), fi = f.imag();
      double* bir = &b.re[(size_t)i * n];
      double* bii = &b.im[(size_t)i * n];
      const double* bkr = &b.re[(size_t)k * n];
      const double* bki = &b.im[(size_t)k * n];
      for (int j = 0; j < n; j++) {
        bir[j] -= fr * bkr[j] - fi * bki[j];
        bii[j] -= fr * bki[j] + fi * bkr[j];
      }
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    cd piv = A(k, k);
    for (int j = 0; j < n; j++) {
      cd v(b.re[(size_t)k * n + j], b.im[(size_t)k * n + j]);
      v /= piv;
      b.re[(size_t)k * n + j] = v.real();
      b.im[(size_t)k * n + j] = v.imag();
    }
    for (int i = 0; i < k; i++) {
      cd f = A(i, k);
      if (f == cd(0)) continue;
      double fr = f.real(), fi = f.imag();
      for (int j = 0; j < n; j++) {
        double xr = b.re[(size_t)k * n + j], xi = b.im[(size_t)k * n + j];
        b.re[(size_t)i * n + j] -= fr * xr - fi * xi;
        b.im[(size_t)i * n + j] -= fr * xi + fi * xr;
      }
    }
  }
}

// Pade coefficients c_j = (2m-j)! m! / ((2m)! j! (m-j)!), j=0..m
static std::vector<double> pade_coeffs(int m) {
  std::vector<double> c(m + 1);
  for (int j = 0; j <= m; j++) {
    double v = 1.0;
    // (2m-j)!/(2m)! * m!/(m-j)! / j!
    for (int t = 0; t < j; t++) v *= (double)(m - t) / (double)(2 * m - t);
    for (int t = 2; t <= j; t++) v /= t;
    c[j] = v;
  }
  return c;
}

// Higham (2005) scaling and squaring with Pade degrees 3,5,7,9,13.
static ZMat expm_pade(const ZMat& a0) {
  static const double theta[5] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1,
                                  2.097847961257068e0, 5.371920351148152e0};
  static const int degs[5] = {3, 5, 7, 9, 13};
  int n = a0.n;
  double nrm = norm1(a0);
  ZMat a = a0;
  int s = 0, m = 13;
  bool found = false;
  for (int t = 0; t < 4; t++)
    if (nrm <= theta[t]) {
      m = degs[t];
      found = true;
      break;
    }
  if (!found) {
    m = 13;
    s = std::max(0, (int)std::ceil(std::log2(nrm / theta[4])));
    if (s > 0) scale(a, std::ldexp(1.0, -s));
  }
  std::vector<double> c = pade_coeffs(m);
  ZMat I = ZMat::eye(n);
  ZMat U(n), V(n);
  ZMat a2 = matmul(a, a);
  if (m < 13) {
    // powers A^2, A^4, ...
    std::vector<ZMat> pw;
    pw.push_back(I);
    pw.push_back(a2);
    for (int k = 2; 2 * k <= m; k++) pw.push_back(matmul(pw.back(), a2));
    ZMat uo(n);
    for (int k = 0; 2 * k + 1 <= m; k++) axpy(uo, c[2 * k + 1], pw[k]);
    for (int k = 0; 2 * k <= m; k++) axpy(V, c[2 * k], pw[k]);
    U = matmul(a, uo);
  } else {
    ZMat a4 = matmul(a2, a2), a6 = matmul(a4, a2);
    ZMat t1(n), t2(n);
    axpy(t1, c[13], a6);
    axpy(t1, c[11], a4);
    axpy(t1, c[9], a2);
    ZMat t1b = matmul(a6, t1);
    axpy(t1b, c[7], a6);
    axpy(t1b, c[5], a4);
    axpy(t1b, c[3], a2);
    axpy(t1b, c[1], I);
    U = matmul(a, t1b);
    axpy(t2, c[12], a6);
    axpy(t2, c[10], a4);
    axpy(t2, c[8], a2);
    V = matmul(a6, t2);
    axpy(V, c[6], a6);
    axpy(V, c[4], a4);
    axpy(V, c[2], a2);
    axpy(V, c[0], I);
  }
  ZMat num = V, den = V;
  axpy(num, 1.0, U);
  axpy(den, -1.0, U);
  solve(den, num);  // num <- (V-U)^{-1}(V+U)
  for (int k = 0; k < s; k++) num = matmul(num, num);
  return num;
}

// clean_up(): tol*round(A/tol) on real and imaginary parts (clean_up.m:70).
static inline double q(double x, double tol) { return tol * std::round((1.0 / tol) * x); }
static void clean(ZMat& a, double tol) {
  if (tol == 0.0) return;
  for (size_t i = 0; i < a.re.size(); i++) {
    a.re[i] = q(a.re[i], tol);
    a.im[i] = q(a.im[i], tol);
  }
}
static bool any_nonzero(const ZMat& a) {
  for (size_t i = 0; i < a.re.size(); i++)
    if (a.re[i] != 0.0 || a.im[i] != 0.0) return true;
  return false;
}

// propagator(L,dt) (propagator.m:34-223) for the dense matrix L.
static ZMat propagator(const ZMat& L, double dt, double small_matrix, double prop_chop, int exact) {
  int n = L.n;
  ZMat A(n);  // -1i*dt*L
  for (size_t i = 0; i < A.re.size(); i++) {
    A.re[i] = dt * L.im[i];
    A.im[i] = -dt * L.re[i];
  }
  if (exact || n < small_matrix) return expm_pade(A);
  clean(A, prop_chop);
  double nrm = norm1(A);
  if (nrm > 1e9) throw std::runtime_error("norm of -i*L*dt exceeds 1e9, check your L and your dt.");
  double chop = (nrm > 1024) ? 2.220446049250313e-16 : prop_chop;
  int nsq = std::max(0, (int)std::ceil(std::log2(nrm)));
  if (nsq > 0) scale(A, 1.0 / std::ldexp(1.0, nsq));
  ZMat P = ZMat::eye(n), next = ZMat::eye(n);
  for (int k = 1; any_nonzero(next); k++) {
    next = matmul(A, next);
    scale(next, 1.0 / k);
    clean(next, chop);
    axpy(P, 1.0, next);
  }
  clean(P, chop);
  for (int k = 0; k < nsq; k++) {
    P = matmul(P, P);
    clean(P, chop);
  }
  return P;
}

// ----------------------------------------------------------------- sparse ops
static void spmv(const Csr& m, const cd* x, cd* y) {
  for (i64 i = 0; i < m.n; i++) {
    cd acc(0);
    for (i64 p = m.ptr[i]; p < m.ptr[i + 1]; p++) acc += m.val[p] * x[m.idx[p]];
    y[i] = acc;
  }
}

static double norm1(const Csr& m) {
  std::vector<double> col(m.n, 0.0);
  for (i64 i = 0; i < m.n; i++)
    for (i64 p = m.ptr[i]; p < m.ptr[i + 1]; p++) col[m.idx[p]] += std::abs(m.val[p]);
  double r = 0;
  for (double v : col) r = std::max(r, v);
  return r;
}

// Planted mistakes, selected by the last entry of the tolerance vector. They
// exist only so that the tests can show that each check of the oracle fails
// on a plausible mistake (0 = none, the faithful restatement).
enum Plant {
  kPlantNone = 0,
  kPlantStepSign = 1,    // step(): the Taylor term with +i instead of -i
  kPlantStepTol = 2,     // step(): the series stopped at 1e-8 instead of eps (dropped terms)
  kPlantDropBin = 3,     // path_trace(): the last bin of binpack() lost
  kPlantDupBox = 4,      // binpack(): each bin after the first also gets the states of the bin before
  kPlantKrylovShift = 5  // krylov(): detection after the step instead of before it
};
static int g_plant = kPlantNone;

// step(L,rho,dt) Liouville branch with reordered Taylor (step.m:204-267,328-375)
static void step_liouv(const Csr& L, std::vector<cd>& rho, double dt) {
  double scaling = 0;
  for (auto& v : rho) scaling = std::max(scaling, std::abs(v));
  if (scaling == 0) return;
  for (auto& v : rho) v /= scaling;
  double norm_mat = norm1(L) * std::abs(dt);
  i64 nsteps = (i64)std::ceil(norm_mat / 2);
  if (nsteps > 10000) throw std::runtime_error("either dt is too long, or L is too big: |L*dt|>1e4");
  if (nsteps < 1) nsteps = 1;  // ceil(0/2)=0 substeps leaves rho unchanged in the reference
  if (norm_mat == 0) nsteps = 0;
  const double tol = (g_plant == kPlantStepTol) ? 1e-8 : 2.220446049250313e-16;
  const double isign = (g_plant == kPlantStepSign) ? 1.0 : -1.0;
  std::vector<cd> next(rho.size()), tmp(rho.size());
  for (i64 s = 0; s < nsteps; s++) {
    next = rho;
    for (int k = 1;; k++) {
      bool any = false;
      for (auto& v : next)
        if (std::abs(v) > tol) {
          any = true;
          break;
        }
      if (!any) break;
      spmv(L, next.data(), tmp.data());
      cd f = cd(0, isign / k) * (dt / (double)nsteps);
      for (size_t i = 0; i < next.size(); i++) {
        next[i] = f * tmp[i];
        rho[i] += next[i];
      }
    }
  }
  for (auto& v : rho) v *= scaling;
}

// Wigner D matrices with rows/columns in descending m (wigner.m:1-3):
// D(m',m) = exp(-i m' alp) d(m',m)(bet) exp(-i m gam). Small-d elements from
// Brink & Satchler's closed forms for l=1,2.
static void wigner_d_small(int l, double b, double* d) {
  double c = std::cos(b), s = std::sin(b);
  if (l == 1) {
    // index 0:+1 1:0 2:-1
    double t[3][3] = {{(1 + c) / 2, -s / std::sqrt(2.0), (1 - c) / 2},
                      {s / std::sqrt(2.0), c, -s / std::sqrt(2.0)},
                      {(1 - c) / 2, s / std::sqrt(2.0), (1 + c) / 2}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) d[i * 3 + j] = t[i][j];
    return;
  }
  double ch = std::cos(b / 2), sh = std::sin(b / 2);
  double r38 = std::sqrt(3.0 / 8.0), r32 = std::sqrt(1.5);
  // upper-left generating block: d(m',m) for m' = 2,1,0
  double d22 = ch * ch * ch * ch, d21 = -s * (1 + c) / 2, d20 = r38 * s * s, d2m1 = -s * (1 - c) / 2,
         d2m2 = sh * sh * sh * sh;
  double d11 = (2 * c - 1) * (1 + c) / 2, d10 = -r32 * s * c, d1m1 = (2 * c + 1) * (1 - c) / 2;
  double d00 = (3 * c * c - 1) / 2;
  int ms[5] = {2, 1, 0, -1, -2};
  auto base = [&](int mp, int m) -> double {
    // reduce to mp>=|m| region with d(m',m) = (-1)^(m-m') d(m,m') = d(-m,-m')
    if (mp == 2) {
      switch (m) {
        case 2: return d22;
        case 1: return d21;
        case 0: return d20;
        case -1: return d2m1;
        default: return d2m2;
      }
    }
    if (mp == 1) {
      switch (m) {
        case 1: return d11;
        case 0: return d10;
        default: return d1m1;
      }
    }
    return d00;
  };
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      int mp = ms[i], m = ms[j];
      double v;
      if (mp >= std::abs(m)) {
        v = base(mp, m);
      } else if (-mp >= std::abs(m)) {
        v = base(-mp, -m) * (((mp - m) % 2 == 0) ? 1.0 : -1.0);  // d(m',m)=(-1)^(m'-m) d(-m',-m)
      } else if (m >= std::abs(mp)) {
        v = base(m, mp) * (((m - mp) % 2 == 0) ? 1.0 : -1.0);  // d(m',m)=(-1)^(m-m') d(m,m')
      } else {
        v = base(-m, -mp);  // d(m',m) = d(-m,-m')
      }
      d[i * 5 + j] = v;
    }
}

static void wigner(int l, double a, double b, double g, cd* D) {
  int sz = 2 * l + 1;
  std::vector<double> d(sz * sz);
  wigner_d_small(l, b, d.data());
  for (int i = 0; i < sz; i++)
    for (int j = 0; j < sz; j++) {
      int mp = l - i, m = l - j;
      D[i * sz + j] = std::exp(cd(0, -mp * a)) * d[i * sz + j] * std::exp(cd(0, -m * g));
    }
}

struct Problem {
  i64 dim;
  std::vector<const Csr*> mats;  // 0 = I, 1..34 = Q slots (nullptr if empty)
  // union pattern of all matrices plus its transpose
  Csr pat;
  std::vector<std::vector<i64>> slot;  // per matrix: position of each nnz in pat
  std::vector<i64> tpos;               // pat position of the transposed element
};

static void build_pattern(Problem& pr) {
  i64 n = pr.dim;
  std::vector<std::vector<i64>> rows(n);
  for (auto* m : pr.mats) {
    if (!m) continue;
    for (i64 i = 0; i < n; i++)
      for (i64 p = m->ptr[i]; p < m->ptr[i + 1]; p++) {
        rows[i].push_back(m->idx[p]);
        rows[m->idx[p]].push_back(i);
      }
  }
  pr.pat.n = n;
  pr.pat.ptr.assign(n + 1, 0);
  for (i64 i = 0; i < n; i++) {
    rows[i].push_back(i);  // keep the diagonal present
    std::sort(rows[i].begin(), rows[i].end());
    rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
    pr.pat.ptr[i + 1] = pr.pat.ptr[i] + (i64)rows[i].size();
  }
  pr.pat.idx.resize(pr.pat.ptr[n]);
  for (i64 i = 0; i < n; i++) std::copy(rows[i].begin(), rows[i].end(), pr.pat.idx.begin() + pr.pat.ptr[i]);
  pr.pat.val.assign(pr.pat.ptr[n], cd(0));
  auto find = [&](i64 i, i64 j) -> i64 {
    auto b = pr.pat.idx.begin() + pr.pat.ptr[i], e = pr.pat.idx.begin() + pr.pat.ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (i64)(it - pr.pat.idx.begin());
  };
  pr.slot.assign(pr.mats.size(), {});
  for (size_t t = 0; t < pr.mats.size(); t++) {
    auto* m = pr.mats[t];
    if (!m) continue;
    pr.slot[t].resize(m->ptr[n]);
    for (i64 i = 0; i < n; i++)
      for (i64 p = m->ptr[i]; p < m->ptr[i + 1]; p++) pr.slot[t][p] = find(i, m->idx[p]);
  }
  pr.tpos.resize(pr.pat.ptr[n]);
  for (i64 i = 0; i < n; i++)
    for (i64 p = pr.pat.ptr[i]; p < pr.pat.ptr[i + 1]; p++) pr.tpos[p] = find(pr.pat.idx[p], i);
}

struct Tols {
  double liouv_zero, zte_tol, prop_chop, small_matrix, krylov_tol, subs_drop, zte_maxden;
  int zte_nsteps, merge_dim;
};

// H(Omega) on the union pattern: (I + orientation(Q,Omega)), symmetrised as
// in orientation.m:57 and powder.m:235.
static Csr assemble(const Problem& pr, double a, double b, double g) {
  cd D1[9], D2[25];
  wigner(1, a, b, g, D1);
  wigner(2, a, b, g, D2);
  std::vector<cd> h(pr.pat.val.size(), cd(0)), q(pr.pat.val.size(), cd(0));
  for (size_t t = 1; t < pr.mats.size(); t++) {
    auto* m = pr.mats[t];
    if (!m) continue;
    int s = (int)t - 1;
    cd c = (s < 9) ? D1[s] : D2[s - 9];
    for (i64 p = 0; p < (i64)m->val.size(); p++) q[pr.slot[t][p]] += c * m->val[p];
  }
  // orientation(): (Hq + Hq')/2
  std::vector<cd> qs(q.size());
  for (size_t p = 0; p < q.size(); p++) qs[p] = (q[p] + std::conj(q[pr.tpos[p]])) * 0.5;
  if (pr.mats[0]) {
    auto* m = pr.mats[0];
    for (i64 p = 0; p < (i64)m->val.size(); p++) h[pr.slot[0][p]] += m->val[p];
  }
  Csr H = pr.pat;
  for (size_t p = 0; p < h.size(); p++) h[p] += qs[p];
  for (size_t p = 0; p < h.size(); p++) H.val[p] = (h[p] + std::conj(h[pr.tpos[p]])) * 0.5;
  return H;
}

static Csr ctranspose(const Csr& m, const std::vector<i64>& tpos) {
  Csr t = m;  // symmetric pattern: L'[i,j] = conj(L[j,i])
  for (size_t p = 0; p < m.val.size(); p++) t.val[p] = std::conj(m.val[tpos[p]]);
  return t;
}

// zte(L,rho) -> kept state indices (zte.m:43-142)
static std::vector<i64> zte(const Csr& L, const std::vector<cd>& rho, const Tols& tl) {
  i64 n = L.n;
  std::vector<i64> all(n);
  std::iota(all.begin(), all.end(), 0);
  i64 nnzr = 0;
  double n1 = 0;
  for (auto& v : rho) {
    if (v != cd(0)) nnzr++;
    n1 += std::abs(v);
  }
  if ((double)nnzr / (double)n > tl.zte_maxden) return all;
  if (n1 < tl.zte_tol) return all;
  double nrm = norm1(L);
  double ts = (nrm == 0) ? 1.0 : 1.0 / nrm;
  std::vector<double> amax(n);
  for (i64 i = 0; i < n; i++) amax[i] = std::abs(rho[i]);
  std::vector<cd> cur = rho;
  auto count = [&]() {
    i64 c = 0;
    for (double v : amax)
      if (v > tl.zte_tol) c++;
    return c;
  };
  for (int s = 2; s <= tl.zte_nsteps; s++) {
    i64 prev = count();
    step_liouv(L, cur, ts);
    for (i64 i = 0; i < n; i++) amax[i] = std::max(amax[i], std::abs(cur[i]));
    if (count() == prev) break;
  }
  std::vector<i64> keep;
  for (i64 i = 0; i < n; i++)
    if (!(amax[i] < tl.zte_tol)) keep.push_back(i);
  return keep;
}

static Csr restrict_(const Csr& m, const std::vector<i64>& states) {
  std::vector<i64> map(m.n, -1);
  for (size_t k = 0; k < states.size(); k++) map[states[k]] = (i64)k;
  Csr r;
  r.n = (i64)states.size();
  r.ptr.assign(r.n + 1, 0);
  for (size_t k = 0; k < states.size(); k++) {
    i64 i = states[k];
    for (i64 p = m.ptr[i]; p < m.ptr[i + 1]; p++)
      if (map[m.idx[p]] >= 0) {
        r.idx.push_back(map[m.idx[p]]);
        r.val.push_back(m.val[p]);
      }
    r.ptr[k + 1] = (i64)r.idx.size();
  }
  return r;
}

// path_trace (path_trace.m:36-183) on a (reduced) generator; returns subspaces
// as lists of state indices, already bin-packed (binpack.m).
static std::vector<std::vector<i64>> path_trace(const Csr& L, const std::vector<cd>& rho, const Tols& tl) {
  i64 n = L.n;
  if (n < tl.merge_dim) {
    std::vector<i64> all(n);
    std::iota(all.begin(), all.end(), 0);
    return {all};
  }
  std::vector<i64> par(n);
  std::iota(par.begin(), par.end(), 0);
  std::function<i64(i64)> root = [&](i64 x) {
    while (par[x] != x) x = par[x] = par[par[x]];
    return x;
  };
  for (i64 i = 0; i < n; i++)
    for (i64 p = L.ptr[i]; p < L.ptr[i + 1]; p++)
      if (std::abs(L.val[p]) > tl.liouv_zero) {
        i64 a = root(i), b = root(L.idx[p]);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
      }
  // component labels in order of their smallest member (scomponents order)
  std::vector<i64> label(n, -1), lab_of_root(n, -1);
  i64 nl = 0;
  for (i64 i = 0; i < n; i++) {
    i64 r = root(i);
    if (lab_of_root[r] < 0) lab_of_root[r] = nl++;
    label[i] = lab_of_root[r];
  }
  std::vector<double> pop(nl, 0.0);
  for (i64 i = 0; i < n; i++) pop[label[i]] += std::abs(rho[i]);
  std::vector<std::vector<i64>> subs;
  for (i64 c = 0; c < nl; c++) {
    if (!(pop[c] > tl.subs_drop)) continue;
    std::vector<i64> st;
    for (i64 i = 0; i < n; i++)
      if (label[i] == c) st.push_back(i);
    subs.push_back(st);
  }
  // binpack: oversize boxes alone, then greedy prefix bins
  std::vector<std::vector<i64>> bins;
  std::vector<size_t> rest;
  for (size_t k = 0; k < subs.size(); k++) {
    if ((i64)subs[k].size() > tl.merge_dim)
      bins.push_back(subs[k]);
    else
      rest.push_back(k);
  }
  while (!rest.empty()) {
    std::vector<i64> bin;
    std::vector<size_t> left;
    i64 acc = 0;
    bool open = true;
    for (size_t k : rest) {
      acc += (i64)subs[k].size();
      if (open && acc <= tl.merge_dim) {
        bin.insert(bin.end(), subs[k].begin(), subs[k].end());
      } else {
        open = false;
        left.push_back(k);
      }
    }
    bins.push_back(bin);
    rest = left;
  }
  if (g_plant == kPlantDropBin && bins.size() > 1) bins.pop_back();
  if (g_plant == kPlantDupBox)
    for (size_t k = bins.size(); k-- > 1;) bins[k].insert(bins[k].end(), bins[k - 1].begin(), bins[k - 1].end());
  return bins;
}

static ZMat dense_of(const Csr& m) {
  ZMat d((int)m.n);
  for (i64 i = 0; i < m.n; i++)
    for (i64 p = m.ptr[i]; p < m.ptr[i + 1]; p++) {
      d.re[(size_t)i * m.n + m.idx[p]] += m.val[p].real();
      d.im[(size_t)i * m.n + m.idx[p]] += m.val[p].imag();
    }
  return d;
}

// evolution(...,'observable') on one subspace (evolution.m:509-576)
static void evolve_sub(const Csr& Lsub, std::vector<cd> rho, std::vector<cd> coil, double dt, i64 nsteps,
                       const Tols& tl, int variant, cd* out) {
  int n = (int)Lsub.n;
  Csr L = Lsub;
  for (auto& v : L.val) v = cd(q(v.real(), tl.liouv_zero), q(v.imag(), tl.liouv_zero));
  for (auto& v : rho) v = cd(q(v.real(), tl.zte_tol), q(v.imag(), tl.zte_tol));
  for (auto& v : coil) v = cd(q(v.real(), tl.zte_tol), q(v.imag(), tl.zte_tol));
  if (n < tl.krylov_tol) {
    ZMat P = propagator(dense_of(L), dt, tl.small_matrix, tl.prop_chop, variant == 1);
    std::vector<double> rr(n), ri(n), nr(n), ni(n), cr(n), ci(n);
    for (int i = 0; i < n; i++) {
      rr[i] = rho[i].real();
      ri[i] = rho[i].imag();
      cr[i] = coil[i].real();
      ci[i] = coil[i].imag();
    }
    for (i64 t = 0; t <= nsteps; t++) {
      double sr = 0, si = 0;  // coil' * rho
      for (int i = 0; i < n; i++) {
        sr += cr[i] * rr[i] + ci[i] * ri[i];
        si += cr[i] * ri[i] - ci[i] * rr[i];
      }
      out[t] += cd(sr, si);
      if (t == nsteps) break;
      for (int i = 0; i < n; i++) {
        const double* pr = &P.re[(size_t)i * n];
        const double* pi = &P.im[(size_t)i * n];
        double ar = 0, ai = 0;
        for (int j = 0; j < n; j++) {
          ar += pr[j] * rr[j] - pi[j] * ri[j];
          ai += pr[j] * ri[j] + pi[j] * rr[j];
        }
        nr[i] = ar;
        ni[i] = ai;
      }
      std::swap(rr, nr);
      std::swap(ri, ni);
    }
  } else {
    // krylov(...,'observable') (krylov.m:139-163)
    for (i64 t = 0; t <= nsteps; t++) {
      if (g_plant == kPlantKrylovShift) step_liouv(L, rho, dt);
      cd s(0);
      for (int i = 0; i < n; i++) s += std::conj(coil[i]) * rho[i];
      out[t] += s;
      if (t == nsteps) break;
      if (g_plant != kPlantKrylovShift) step_liouv(L, rho, dt);
    }
  }
}

static Csr from_arrays(i64 n, const i64* ptr, const i64* idx, const double* val) {
  Csr m;
  m.n = n;
  m.ptr.assign(ptr, ptr + n + 1);
  i64 nnz = ptr[n];
  m.idx.assign(idx, idx + nnz);
  m.val.resize(nnz);
  for (i64 p = 0; p < nnz; p++) m.val[p] = cd(val[2 * p], val[2 * p + 1]);
  return m;
}

// Tolerance vector (kernel/utilities/tolerances.m names): liouv_zero, zte_tol,
// prop_chop, small_matrix, krylov_tol, subs_drop, zte_maxden, zte_nsteps,
// merge_dim, then the planted-mistake code (0 in every use but the tests);
// refcpu_powder reads one more entry, the acquisition's dead time
// (parameters.dead_time, 0 = none).
static Tols read_tols(const double* tols) {
  Tols tl;
  tl.liouv_zero = tols[0];
  tl.zte_tol = tols[1];
  tl.prop_chop = tols[2];
  tl.small_matrix = tols[3];
  tl.krylov_tol = tols[4];
  tl.subs_drop = tols[5];
  tl.zte_maxden = tols[6];
  tl.zte_nsteps = (int)tols[7];
  tl.merge_dim = (int)tols[8];
  g_plant = (int)tols[9];
  return tl;
}

static Problem make_problem(i64 dim, const i64* const* ptrs, const i64* const* idxs, const double* const* vals,
                            std::vector<Csr>& store) {
  Problem pr;
  pr.dim = dim;
  pr.mats.assign(35, nullptr);
  for (int t = 0; t < 35; t++)
    if (ptrs[t]) {
      store[t] = from_arrays(dim, ptrs[t], idxs[t], vals[t]);
      pr.mats[t] = &store[t];
    }
  build_pattern(pr);
  return pr;
}

// The per-orientation reduction of evolution.m:141-145 (reduce(L',coil):
// zte() then path_trace() with binpack()): the state lists of the subspaces,
// in the order the reference propagates them.
static std::vector<std::vector<i64>> reduce_orientation(const Problem& pr, double a, double b, double g,
                                                        const std::vector<cd>& c0, const Tols& tl) {
  Csr H = assemble(pr, a, b, g);
  Csr Lt = ctranspose(H, pr.tpos);
  std::vector<i64> kz = zte(Lt, c0, tl);
  Csr Lz = restrict_(Lt, kz);
  std::vector<cd> cz(kz.size());
  for (size_t k = 0; k < kz.size(); k++) cz[k] = c0[kz[k]];
  std::vector<std::vector<i64>> subs;
  for (auto& s : path_trace(Lz, cz, tl)) {
    std::vector<i64> glob(s.size());
    for (size_t k = 0; k < s.size(); k++) glob[k] = kz[s[k]];
    subs.push_back(glob);
  }
  return subs;
}

}  // namespace

extern "C" {

// Thread-local description of the last failure.
const char* refcpu_last_error() { return g_err.c_str(); }

// Reduction variants: 0 = faithful per-orientation reduce() (ZTE + path
// tracing) with the reference's propagator branch choice (expm below
// small_matrix, chop-Taylor above); 1 = same reduction, exact Pade expm for
// every subspace size; 2 = single union-pattern reduction (coil-connected
// component of the union sparsity pattern), exact expm.
//
// mats: 35 CSR matrices, slot 0 = I, slots 1..34 = Q in rank-1 (k,m)
// row-major then rank-2 (k,m) row-major order; a slot with ptr == NULL is
// empty. Complex values are interleaved (re, im).
int refcpu_powder(i64 dim, const i64* const* ptrs, const i64* const* idxs, const double* const* vals,
                  const double* alphas, const double* betas, const double* gammas, const double* weights,
                  i64 n_or, const double* rho0, const double* coil, double dt, i64 nsteps, const double* tols,
                  int variant, int nthreads, double* fid_out, double* per_orient_out) {
  try {
    Tols tl = read_tols(tols);
    const double dead_time = tols[10];
    std::vector<Csr> store(35);
    Problem pr = make_problem(dim, ptrs, idxs, vals, store);
    std::vector<cd> r0(dim), c0(dim);
    for (i64 i = 0; i < dim; i++) {
      r0[i] = cd(rho0[2 * i], rho0[2 * i + 1]);
      c0[i] = cd(coil[2 * i], coil[2 * i + 1]);
    }
    // union-pattern reduction for variant 2: coil-connected component
    std::vector<i64> ustates;
    if (variant == 2) {
      std::vector<i64> par(dim);
      std::iota(par.begin(), par.end(), 0);
      auto root = [&](i64 x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
      };
      for (i64 i = 0; i < dim; i++)
        for (i64 p = pr.pat.ptr[i]; p < pr.pat.ptr[i + 1]; p++) {
          i64 a = root(i), b = root(pr.pat.idx[p]);
          if (a != b) par[std::max(a, b)] = std::min(a, b);
        }
      std::vector<char> hit(dim, 0);
      for (i64 i = 0; i < dim; i++)
        if (c0[i] != cd(0)) hit[root(i)] = 1;
      for (i64 i = 0; i < dim; i++)
        if (hit[root(i)]) ustates.push_back(i);
    }
    i64 npts = nsteps + 1;
    std::vector<cd> per((size_t)n_or * npts, cd(0));
    std::string err;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (i64 o = 0; o < n_or; o++) {
      try {
        Csr H = assemble(pr, alphas[o], betas[o], gammas[o]);
        cd* out = &per[(size_t)o * npts];
        // dead time: rho0 = step(L, rho0, dead_time) on the whole generator (acquire.m:65-67)
        std::vector<cd> r0o = r0;
        if (dead_time > 0) step_liouv(H, r0o, dead_time);
        std::vector<std::vector<i64>> subs;
        if (variant == 2)
          subs.push_back(ustates);
        else
          subs = reduce_orientation(pr, alphas[o], betas[o], gammas[o], c0, tl);
        for (auto& s : subs) {
          Csr Ls = restrict_(H, s);
          std::vector<cd> rs(s.size()), cs(s.size());
          for (size_t k = 0; k < s.size(); k++) {
            rs[k] = r0o[s[k]];
            cs[k] = c0[s[k]];
          }
          evolve_sub(Ls, rs, cs, dt, nsteps, tl, variant == 0 ? 0 : 1, out);
        }
      } catch (std::exception& e) {
#ifdef _OPENMP
#pragma omp critical
#endif
        err = e.what();
      }
    }
    if (!err.empty()) throw std::runtime_error(err);
    // weighted sum in grid order (powder.m:280-286)
    for (i64 t = 0; t < npts; t++) {
      cd acc = weights[0] * per[t];
      for (i64 o = 1; o < n_or; o++) acc += weights[o] * per[(size_t)o * npts + t];
      fid_out[2 * t] = acc.real();
      fid_out[2 * t + 1] = acc.imag();
    }
    g_plant = kPlantNone;
    if (per_orient_out)
      for (size_t k = 0; k < per.size(); k++) {
        per_orient_out[2 * k] = per[k].real();
        per_orient_out[2 * k + 1] = per[k].imag();
      }
    return 0;
  } catch (std::exception& e) {
    g_plant = kPlantNone;
    g_err = e.what();
    return 1;
  }
}

// Dense helpers exposed for the oracle's own tests.
int refcpu_expm(int n, const double* a, double* out) {
  try {
    ZMat m(n);
    for (int i = 0; i < n * n; i++) {
      m.re[i] = a[2 * i];
      m.im[i] = a[2 * i + 1];
    }
    ZMat e = expm_pade(m);
    for (int i = 0; i < n * n; i++) {
      out[2 * i] = e.re[i];
      out[2 * i + 1] = e.im[i];
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void refcpu_wigner(int l, double a, double b, double g, double* out) {
  int sz = 2 * l + 1;
  std::vector<cd> D(sz * sz);
  wigner(l, a, b, g, D.data());
  for (int i = 0; i < sz * sz; i++) {
    out[2 * i] = D[i].real();
    out[2 * i + 1] = D[i].imag();
  }
}

// propagator() on a dense generator with the reference's branch choice.
int refcpu_propagator(int n, const double* L, double dt, double small_matrix, double prop_chop, double* out) {
  try {
    ZMat m(n);
    for (int i = 0; i < n * n; i++) {
      m.re[i] = L[2 * i];
      m.im[i] = L[2 * i + 1];
    }
    ZMat P = propagator(m, dt, small_matrix, prop_chop, 0);
    for (int i = 0; i < n * n; i++) {
      out[2 * i] = P.re[i];
      out[2 * i + 1] = P.im[i];
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// step(L,rho,dt) (kernel/step.m:204-267, reordered_taylor :328-375) for one
// CSR generator; rho (n interleaved complex) is overwritten. plant selects a
// planted mistake (0 = none).
int refcpu_step(i64 n, const i64* ptr, const i64* idx, const double* val, double* rho, double dt, int plant) {
  try {
    Csr L = from_arrays(n, ptr, idx, val);
    std::vector<cd> r(n);
    for (i64 i = 0; i < n; i++) r[i] = cd(rho[2 * i], rho[2 * i + 1]);
    g_plant = plant;
    step_liouv(L, r, dt);
    g_plant = kPlantNone;
    for (i64 i = 0; i < n; i++) {
      rho[2 * i] = r[i].real();
      rho[2 * i + 1] = r[i].imag();
    }
    return 0;
  } catch (std::exception& e) {
    g_plant = kPlantNone;
    g_err = e.what();
    return 1;
  }
}

// The subspaces reduce(L',coil) keeps at one orientation: their count is
// returned (negative on failure) and their dimensions written to dims_out
// (at most max_out entries); states_out (optional, at most dim entries)
// receives the state indices of the subspaces one after another.
i64 refcpu_reduce(i64 dim, const i64* const* ptrs, const i64* const* idxs, const double* const* vals, double alpha,
                  double beta, double gamma, const double* coil, const double* tols, i64* dims_out, i64 max_out,
                  i64* states_out) {
  try {
    Tols tl = read_tols(tols);
    std::vector<Csr> store(35);
    Problem pr = make_problem(dim, ptrs, idxs, vals, store);
    std::vector<cd> c0(dim);
    for (i64 i = 0; i < dim; i++) c0[i] = cd(coil[2 * i], coil[2 * i + 1]);
    auto subs = reduce_orientation(pr, alpha, beta, gamma, c0, tl);
    g_plant = kPlantNone;
    i64 pos = 0;
    for (size_t k = 0; k < subs.size(); k++) {
      if ((i64)k < max_out) dims_out[k] = (i64)subs[k].size();
      if (states_out)
        for (i64 v : subs[k])
          if (pos < dim) states_out[pos++] = v;
    }
    return (i64)subs.size();
  } catch (std::exception& e) {
    g_plant = kPlantNone;
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"

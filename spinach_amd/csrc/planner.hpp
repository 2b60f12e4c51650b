// Host-side planning of libspinach_powder (plain C++17, no device code, so
// that the CPU sanitizer build in tests/native compiles it with g++):
// reading MATLAB CSC input, the coil-connected reduction on the union
// sparsity pattern, the real-generator test, tolerances, and the sliced-ELL
// layout of the stepping path.
//
// The reduction replaces the per-orientation reduce(L',coil) of
// kernel/evolution.m:141-145: the coil-connected component of the union
// pattern of I and every Q slot is an invariant subspace of every L(Omega)
// that contains every state the coil can see, so restricting to it is exact
// (kernel/utilities/zte.m and path_trace.m keep the same component up to
// their drop thresholds).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/spinach_powder.h"

namespace spp {

struct SpError : std::runtime_error {
  int code;
  SpError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Triplets {
  std::vector<int64_t> row, col;
  std::vector<double> re, im;
};

inline Triplets read_csc(const sp_csc* m, int64_t dim, const char* what) {
  Triplets t;
  if (!m || m->n == 0) return t;
  if (m->n != dim) throw SpError(SP_EINVAL, std::string(what) + ": dimension does not match I");
  if (!m->jc) throw SpError(SP_EINVAL, std::string(what) + ": NULL column pointers");
  uint64_t nnz = m->jc[dim];
  if (nnz == 0) return t;
  if (!m->ir || !m->pr) throw SpError(SP_EINVAL, std::string(what) + ": NULL row indices or values");
  if (m->jc[0] != 0) throw SpError(SP_EINVAL, std::string(what) + ": jc[0] must be 0");
  for (int64_t c = 0; c < dim; c++) {
    if (m->jc[c + 1] < m->jc[c] || m->jc[c + 1] > nnz)
      throw SpError(SP_EINVAL, std::string(what) + ": column pointers decrease or exceed jc[n]");
    for (uint64_t p = m->jc[c]; p < m->jc[c + 1]; p++) {
      uint64_t r = m->ir[p];
      if (r >= (uint64_t)dim) throw SpError(SP_EINVAL, std::string(what) + ": row index out of range");
      double re = m->is_complex ? m->pr[2 * p] : m->pr[p];
      double im = m->is_complex ? m->pr[2 * p + 1] : 0.0;
      if (!std::isfinite(re) || !std::isfinite(im))
        throw SpError(SP_EINVAL, std::string(what) + ": non-finite value");
      if (re == 0.0 && im == 0.0) continue;
      t.row.push_back((int64_t)r);
      t.col.push_back(c);
      t.re.push_back(re);
      t.im.push_back(im);
    }
  }
  return t;
}

// spin_system.tols fields the library reads (kernel/utilities/tolerances.m defaults).
struct Tols {
  double liouv_zero = 1e-10, zte_tol = 1e-24, prop_chop = 1e-10, small_matrix = 200, krylov_tol = 10000;
};

inline Tols read_tols(const sp_tols* t) {
  Tols r;
  if (!t) return r;
  const double v[5] = {t->liouv_zero, t->zte_tol, t->prop_chop, t->small_matrix, t->krylov_tol};
  for (double x : v)
    if (!std::isfinite(x) || x < 0) throw SpError(SP_EINVAL, "tolerances must be finite and non-negative");
  if (t->krylov_tol < 1) throw SpError(SP_EINVAL, "krylov_tol must be at least 1");
  r.liouv_zero = t->liouv_zero;
  r.zte_tol = t->zte_tol;
  r.prop_chop = t->prop_chop;
  r.small_matrix = t->small_matrix;
  r.krylov_tol = t->krylov_tol;
  return r;
}

// Largest coil-connected dimension of the dense paths, and of the stepping
// path (four state vectors of the padded dimension in LDS, 64 B per row).
constexpr int32_t kDenseMaxDim = 1024;
constexpr int32_t kStepMaxDim = 2304;

// Reduced, device-ready description of one spin system.
struct SysPlan {
  int64_t full_dim = 0;
  int32_t n = 0, np = 0, nnz = 0;
  std::vector<int32_t> row, col, tpos, slots;
  std::vector<double> vals;  // interleaved complex, (1+nq)*nnz
  std::vector<double> rho, coil;
  std::vector<double> alphas, betas, gammas, weights;
  bool real = false;  // generator real at every orientation
  // relaxation superoperator (sp_relax): R0 on the union pattern (interleaved complex,
  // nnz entries; empty without R0) and the seven diagonal planes on the reduced
  // states (7*n; empty without them)
  bool relax = false;
  std::vector<double> rvals, rdiag;
  std::vector<int64_t> states;  // full-space index of every reduced state
};

// Partner slot and sign of the conjugation symmetry of the Wigner matrices:
// conj(D^r(k,m)) = (-1)^(k-m) D^r(-k,-m) (slot order: descending k and m).
inline void conj_partner(int t, int& partner, double& sign) {
  if (t < 9) {
    const int i = t / 3, j = t % 3;
    partner = (2 - i) * 3 + (2 - j);
    sign = ((i - j) & 1) ? -1.0 : 1.0;
  } else {
    const int i = (t - 9) / 5, j = (t - 9) % 5;
    partner = 9 + (4 - i) * 5 + (4 - j);
    sign = ((i - j) & 1) ? -1.0 : 1.0;
  }
}

// H(Omega) = I + sum D^r(k,m)(Omega) Q{r}{k,m} is real at every orientation
// exactly when I is real and Q{r}{k,m} = (-1)^(k-m) conj(Q{r}{-k,-m}) for
// every slot (the D^r(k,m) are linearly independent functions of Omega).
// Checked on value planes laid out as plan_system lays them out (I first,
// then the non-empty slots), to 8 ulp of the largest magnitude: a mismatch
// at that level is below the rounding of the complex path itself.
inline bool generator_is_real(const std::vector<int32_t>& slots, const std::vector<double>& vals, size_t nnz) {
  double scale = 0.0;
  for (double v : vals) scale = std::max(scale, std::fabs(v));
  const double tol = 8.0 * 2.220446049250313e-16 * scale;
  for (size_t e = 0; e < nnz; e++)
    if (std::fabs(vals[2 * e + 1]) > tol) return false;
  std::vector<const double*> plane(SP_NQ, nullptr);
  for (size_t q = 0; q < slots.size(); q++) plane[slots[q]] = vals.data() + 2 * (q + 1) * nnz;
  for (int t = 0; t < SP_NQ; t++) {
    int u;
    double sg;
    conj_partner(t, u, sg);
    const double* a = plane[t];
    const double* b = plane[u];
    if (!a && !b) continue;
    for (size_t e = 0; e < nnz; e++) {
      const double ar = a ? a[2 * e] : 0.0, ai = a ? a[2 * e + 1] : 0.0;
      const double br = b ? b[2 * e] : 0.0, bi = b ? b[2 * e + 1] : 0.0;
      if (std::fabs(ar - sg * br) > tol || std::fabs(ai + sg * bi) > tol) return false;
    }
  }
  return true;
}

inline SysPlan plan_system(const sp_csc* I, const sp_csc* Q, const double* al, const double* be, const double* ga,
                    const double* w, int64_t n_or, const double* rho0, const double* coil,
                    const sp_relax* rlx = nullptr) {
  if (!I) throw SpError(SP_EINVAL, "I is NULL");
  if (!Q) throw SpError(SP_EINVAL, "Q is NULL");
  if (!al || !be || !ga || !w) throw SpError(SP_EINVAL, "grid arrays must not be NULL");
  if (n_or <= 0) throw SpError(SP_EINVAL, "the grid must have at least one orientation");
  if (!rho0 || !coil) throw SpError(SP_EINVAL, "rho0 and coil must not be NULL");
  const int64_t dim = I->n;
  if (dim <= 0) throw SpError(SP_EINVAL, "I must have a positive dimension");
  if (dim > (int64_t)INT32_MAX) throw SpError(SP_ERANGE, "dimension exceeds the 32-bit index range");
  std::vector<Triplets> mats(1 + SP_NQ);
  mats[0] = read_csc(I, dim, "I");
  for (int q = 0; q < SP_NQ; q++) {
    char what[32];
    snprintf(what, sizeof(what), "Q slot %d", q);
    mats[1 + q] = read_csc(&Q[q], dim, what);
  }
  for (int64_t k = 0; k < n_or; k++)
    if (!std::isfinite(al[k]) || !std::isfinite(be[k]) || !std::isfinite(ga[k]) || !std::isfinite(w[k]))
      throw SpError(SP_EINVAL, "non-finite grid entry");
  for (int64_t i = 0; i < 2 * dim; i++)
    if (!std::isfinite(rho0[i]) || !std::isfinite(coil[i])) throw SpError(SP_EINVAL, "non-finite state vector");
  // relaxation: R0 joins the union pattern (mats[1 + SP_NQ]) without a value plane of its own in vals
  Triplets r0;
  const bool have_diag = rlx && rlx->diag;
  if (rlx && rlx->R0) r0 = read_csc(rlx->R0, dim, "R0");
  if (have_diag) {
    for (int64_t i = 0; i < 7 * dim; i++)
      if (!std::isfinite(rlx->diag[i])) throw SpError(SP_EINVAL, "non-finite relaxation diagonal");
    // the diagonal joins the pattern (value 0 in R0's plane), so that the stepping path's sparse
    // layout has a place for the per-orientation rates; the reduction is unchanged (it follows
    // off-diagonal entries only)
    for (int64_t i = 0; i < dim; i++) {
      r0.row.push_back(i);
      r0.col.push_back(i);
      r0.re.push_back(0.0);
      r0.im.push_back(0.0);
    }
  }
  mats.push_back(r0);

  // union adjacency (symmetrised) and the coil-connected component
  std::vector<std::vector<int64_t>> adj(dim);
  for (auto& m : mats)
    for (size_t p = 0; p < m.row.size(); p++)
      if (m.row[p] != m.col[p]) {
        adj[m.row[p]].push_back(m.col[p]);
        adj[m.col[p]].push_back(m.row[p]);
      }
  std::vector<char> seen(dim, 0);
  std::vector<int64_t> stack;
  for (int64_t i = 0; i < dim; i++)
    if (coil[2 * i] != 0.0 || coil[2 * i + 1] != 0.0) {
      seen[i] = 1;
      stack.push_back(i);
    }
  while (!stack.empty()) {
    int64_t v = stack.back();
    stack.pop_back();
    for (int64_t u : adj[v])
      if (!seen[u]) {
        seen[u] = 1;
        stack.push_back(u);
      }
  }
  std::vector<int64_t> map(dim, -1);
  SysPlan sp;
  sp.full_dim = dim;
  int32_t n = 0;
  for (int64_t i = 0; i < dim; i++)
    if (seen[i]) {
      map[i] = n++;
      sp.states.push_back(i);
    }
  sp.n = n;
  sp.np = std::max<int32_t>(16, (n + 15) / 16 * 16);
  if (n > kStepMaxDim)
    throw SpError(SP_ERANGE, "the coil-connected subspace has dimension " + std::to_string(n) +
                                 "; the dense paths support up to " + std::to_string(kDenseMaxDim) +
                                 " and the stepping path up to " + std::to_string(kStepMaxDim));
  // reduced union pattern
  std::vector<std::pair<int64_t, int64_t>> ent;
  for (auto& m : mats)
    for (size_t p = 0; p < m.row.size(); p++) {
      int64_t r = map[m.row[p]], c = map[m.col[p]];
      if (r < 0 || c < 0) continue;
      ent.emplace_back(r, c);
      ent.emplace_back(c, r);
    }
  std::sort(ent.begin(), ent.end());
  ent.erase(std::unique(ent.begin(), ent.end()), ent.end());
  sp.nnz = (int32_t)ent.size();
  auto find = [&](int64_t r, int64_t c) -> int64_t {
    auto it = std::lower_bound(ent.begin(), ent.end(), std::make_pair(r, c));
    return (int64_t)(it - ent.begin());
  };
  sp.row.resize(sp.nnz);
  sp.col.resize(sp.nnz);
  sp.tpos.resize(sp.nnz);
  for (int32_t e = 0; e < sp.nnz; e++) {
    sp.row[e] = (int32_t)ent[e].first;
    sp.col[e] = (int32_t)ent[e].second;
    sp.tpos[e] = (int32_t)find(ent[e].second, ent[e].first);
  }
  // value arrays: I first, then every non-empty Q slot
  std::vector<std::vector<double>> planes;
  auto scatter = [&](const Triplets& m, std::vector<double>& v) {
    bool any = false;
    v.assign(2 * (size_t)sp.nnz, 0.0);
    for (size_t p = 0; p < m.row.size(); p++) {
      int64_t r = map[m.row[p]], c = map[m.col[p]];
      if (r < 0 || c < 0) continue;
      int64_t e = find(r, c);
      v[2 * e] += m.re[p];
      v[2 * e + 1] += m.im[p];
      any = true;
    }
    return any;
  };
  std::vector<double> tmp;
  scatter(mats[0], tmp);
  sp.vals = tmp;
  for (int q = 0; q < SP_NQ; q++)
    if (scatter(mats[1 + q], tmp)) {
      sp.slots.push_back(q);
      sp.vals.insert(sp.vals.end(), tmp.begin(), tmp.end());
    }
  if (scatter(mats[1 + SP_NQ], tmp)) sp.rvals = tmp;
  if (have_diag) {
    sp.rdiag.assign(7 * (size_t)n, 0.0);
    for (int64_t i = 0; i < dim; i++)
      if (map[i] >= 0)
        for (int p = 0; p < 7; p++) sp.rdiag[(size_t)p * n + map[i]] = rlx->diag[(size_t)p * dim + i];
  }
  sp.relax = !sp.rvals.empty() || have_diag;
  sp.rho.assign(2 * (size_t)n, 0.0);
  sp.coil.assign(2 * (size_t)n, 0.0);
  for (int64_t i = 0; i < dim; i++)
    if (map[i] >= 0) {
      sp.rho[2 * map[i]] = rho0[2 * i];
      sp.rho[2 * map[i] + 1] = rho0[2 * i + 1];
      sp.coil[2 * map[i]] = coil[2 * i];
      sp.coil[2 * map[i] + 1] = coil[2 * i + 1];
    }
  // L = H + iR is not Hermitian: no real-generator (spectral, real-product) path
  sp.real = !sp.relax && generator_is_real(sp.slots, sp.vals, (size_t)sp.nnz);
  sp.alphas.assign(al, al + n_or);
  sp.betas.assign(be, be + n_or);
  sp.gammas.assign(ga, ga + n_or);
  sp.weights.assign(w, w + n_or);
  return sp;
}

// The realness test on the full (unreduced) matrices, for sp_generator_is_real.
inline bool generator_is_real_full(const sp_csc* I, const sp_csc* Q) {
  if (!I) throw SpError(SP_EINVAL, "I is NULL");
  if (!Q) throw SpError(SP_EINVAL, "Q is NULL");
  const int64_t dim = I->n;
  if (dim <= 0) throw SpError(SP_EINVAL, "I must have a positive dimension");
  std::vector<Triplets> mats(1 + SP_NQ);
  mats[0] = read_csc(I, dim, "I");
  for (int q = 0; q < SP_NQ; q++) {
    char what[32];
    snprintf(what, sizeof(what), "Q slot %d", q);
    mats[1 + q] = read_csc(&Q[q], dim, what);
  }
  std::vector<std::pair<int64_t, int64_t>> ent;
  for (auto& m : mats)
    for (size_t p = 0; p < m.row.size(); p++) ent.emplace_back(m.row[p], m.col[p]);
  std::sort(ent.begin(), ent.end());
  ent.erase(std::unique(ent.begin(), ent.end()), ent.end());
  const size_t nnz = ent.size();
  std::vector<int32_t> slots;
  std::vector<double> vals;
  for (int q = 0; q <= SP_NQ; q++) {
    const Triplets& m = mats[q];
    if (q > 0 && m.row.empty()) continue;
    std::vector<double> v(2 * nnz, 0.0);
    for (size_t p = 0; p < m.row.size(); p++) {
      const size_t e = (size_t)(std::lower_bound(ent.begin(), ent.end(), std::make_pair(m.row[p], m.col[p])) -
                                ent.begin());
      v[2 * e] += m.re[p];
      v[2 * e + 1] += m.im[p];
    }
    if (q > 0) slots.push_back(q - 1);
    vals.insert(vals.end(), v.begin(), v.end());
  }
  return generator_is_real(slots, vals, nnz);
}

// Sliced-ELL layout of the stepping path. Each slice holds 64 / G rows and is
// read by one wave: the G lanes of a row take its entries k = q, q+G, ... so
// that entry k of row r (r' = r - first row of the slice) sits at
//   start[s] + 64 * (k / G) + G * r' + k % G
// and one load instruction of the wave reads 64 consecutive entries. Padding
// entries point at their own row and hold zero.
constexpr int kSellWave = 64;

struct SellDesc {
  int32_t nslices;    // ceil(n / (64 / G))
  int32_t sell_nnz;   // entries of the slice layout, padding included
  int64_t slice_off;  // offset of the system's slice starts in the plan's array
  int64_t sell_off;   // offset of the system's slice-layout columns in the plan's array
  int32_t lanes;      // G: lanes per row (1, 2, 4 or 8)
  int32_t pad;
};

// Lanes per row from the mean row length: enough lanes that a row's entries
// take a few loads each, few enough that a slice still spans many rows.
inline int sell_lanes(int n, int nnz) {
  const double mean = n > 0 ? (double)nnz / n : 0.0;
  return mean >= 16 ? 8 : (mean >= 8 ? 4 : (mean >= 3 ? 2 : 1));
}

// Slices and widest slice (in 64-entry columns) of the layout with G lanes per row.
inline void sell_shape(const SysPlan& sp, int G, int& nslices, int& wmax) {
  const int n = sp.n, H = kSellWave / G;
  std::vector<int32_t> rowlen(n, 0);
  for (int32_t e = 0; e < sp.nnz; e++) rowlen[sp.row[e]]++;
  nslices = (n + H - 1) / H;
  wmax = 0;
  for (int r = 0; r < n; r++) wmax = std::max(wmax, (rowlen[r] + G - 1) / G);
}

// Lanes per row for the register-cached product of the stepping kernel with rcs (1 or 2)
// slices per wave (slice s goes to wave s % 4) of at most entries / rcs columns: the fewest
// lanes per row that fit (no cross-lane sums with one lane per row); 0 when none fits.
inline int sell_lanes_cached(const SysPlan& sp, int rcs, int entries) {
  for (int G = 1; G <= 8; G <<= 1) {
    int ns = 0, w = 0;
    sell_shape(sp, G, ns, w);
    if ((ns + 3) / 4 <= rcs && w <= entries / rcs) return G;
  }
  return 0;
}

// Reverse Cuthill-McKee order of the reduced union pattern (order[k] = the state placed
// at position k): breadth-first from a least-connected state, neighbours by increasing
// degree, reversed. Rows that are neighbours in the order then have their k-th entries
// in neighbouring columns, so a wave's gathers of the state hit neighbouring LDS words.
inline std::vector<int32_t> rcm_order(const SysPlan& sp) {
  const int n = sp.n;
  std::vector<int32_t> first(n + 1, 0), deg(n, 0);
  for (int32_t e = 0; e < sp.nnz; e++) deg[sp.row[e]]++;
  for (int i = 0; i < n; i++) first[i + 1] = first[i] + deg[i];  // the pattern is sorted by row
  std::vector<char> seen(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> nb;
  for (int done = 0; done < n;) {
    int start = -1;
    for (int i = 0; i < n; i++)
      if (!seen[i] && (start < 0 || deg[i] < deg[start])) start = i;
    seen[start] = 1;
    size_t head = order.size();
    order.push_back(start);
    done++;
    while (head < order.size()) {
      const int u = order[head++];
      nb.clear();
      for (int32_t e = first[u]; e < first[u + 1]; e++)
        if (!seen[sp.col[e]]) {
          seen[sp.col[e]] = 1;
          nb.push_back(sp.col[e]);
        }
      std::stable_sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
      for (int32_t v : nb) order.push_back(v);
      done += (int)nb.size();
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Sliced-ELL layout of the reduced union pattern in the state order `order` (position k
// holds state order[k]; empty: the reduced order): row r' of slice s is position s*H + r',
// its entries sorted by position of their column, entry k at start[s] + 64*(k/G) + G*r' +
// k%G; colv holds positions, pos maps each union-pattern entry to its layout slot.
inline void sell_layout(const SysPlan& sp, std::vector<int32_t>& start, std::vector<int32_t>& colv,
                        std::vector<int32_t>& pos, SellDesc& d, int lanes = 0,
                        const std::vector<int32_t>& order = {}) {
  const int n = sp.n;
  const int G = lanes > 0 ? lanes : sell_lanes(sp.n, sp.nnz);
  const int H = kSellWave / G;  // rows per slice
  const int ns = (n + H - 1) / H;
  std::vector<int32_t> at(n);  // position of every state
  for (int i = 0; i < n; i++) at[i] = order.empty() ? i : -1;
  for (int k = 0; k < (int)order.size(); k++) at[order[k]] = k;
  std::vector<int32_t> rowlen(n, 0), first(n + 1, 0);
  for (int32_t e = 0; e < sp.nnz; e++) rowlen[at[sp.row[e]]]++;
  for (int i = 0; i < n; i++) first[i + 1] = first[i] + rowlen[i];
  // the rank of every entry within its row, by the position of its column
  std::vector<int32_t> rank(sp.nnz);
  {
    std::vector<std::vector<std::pair<int32_t, int32_t>>> rows(n);
    for (int32_t e = 0; e < sp.nnz; e++) rows[at[sp.row[e]]].emplace_back(at[sp.col[e]], e);
    for (auto& r : rows) {
      std::sort(r.begin(), r.end());
      for (int32_t k = 0; k < (int32_t)r.size(); k++) rank[r[k].second] = k;
    }
  }
  d.nslices = ns;
  d.lanes = G;
  d.pad = 0;
  d.slice_off = (int64_t)start.size();
  d.sell_off = (int64_t)colv.size();
  std::vector<int32_t> st(ns + 1, 0);
  for (int s2 = 0; s2 < ns; s2++) {
    int w = 0;
    for (int r = s2 * H; r < std::min(n, (s2 + 1) * H); r++) w = std::max(w, (rowlen[r] + G - 1) / G);
    st[s2 + 1] = st[s2] + kSellWave * w;
  }
  d.sell_nnz = st[ns];
  start.insert(start.end(), st.begin(), st.end());
  const size_t base = colv.size();
  colv.resize(base + (size_t)d.sell_nnz);
  for (int s2 = 0; s2 < ns; s2++)
    for (int k = st[s2]; k < st[s2 + 1]; k++) colv[base + k] = s2 * H + ((k - st[s2]) % kSellWave) / G;
  for (int32_t e = 0; e < sp.nnz; e++) {
    const int r = at[sp.row[e]], k = rank[e];
    const int s2 = r / H, rr = r - s2 * H;
    const int p2 = st[s2] + kSellWave * (k / G) + G * rr + k % G;
    colv[base + p2] = at[sp.col[e]];
    pos.push_back(p2);
  }
}

}  // namespace spp

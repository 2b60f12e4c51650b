// Native host setup: the isotropic part I and the 34 anisotropic components Q{r}{k,m} of the
// Liouvillian commutation superoperator in the spin-1/2 sphten-liouv basis, from the
// Hamiltonian's term list (kernel/hamiltonian.m:673-792 with the product-table construction
// of kernel/operators/superop.m:62-170 and kernel/cache/ist_product_table.m:42-125).
//
// Basis states are rows of per-spin IST indices (0 = unit, 1 = T(1,+1), 2 = T(1,0),
// 3 = T(1,-1)), sorted lexicographically, so the base-4 keys of the rows ascend and a state is
// found by binary search. A term is the product of one or two single-spin ISTs with a real
// coefficient (the operator names' expansion, 'L+' = -sqrt(2) T(1,+1), 'L-' = sqrt(2) T(1,-1),
// 'Lz' = T(1,0)); its commutation superoperator is left - right, each side the product over the
// active spins of the structure coefficients c[op][source digit][result digit] with the other
// spins' digits unchanged, results outside the basis dropped. Plain C++ (host code only).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace spb {

typedef std::complex<double> cd;

// Structure coefficients of the spin-1/2 IST algebra, normalised states:
// T_n * (T_m/|T_m|) = sum_k left[n][m][k] T_k/|T_k| (and right with the product reversed).
struct NzList {  // the nonzero result digits d[] and coefficients c[] of one (operator, source digit)
  int n = 0;
  int d[4];
  double c[4];
};

struct IstTables {
  double left[4][4][4], right[4][4][4];
  NzList nz[2][4][4];  // [side][operator][source digit]
  IstTables() {
    // 2 x 2 matrices of E, T(1,+1), T(1,0), T(1,-1) (irr_sph_ten.m:54-75: top state -L+/sqrt(2),
    // then Racah lowering with L-)
    cd t[4][2][2] = {};
    t[0][0][0] = t[0][1][1] = 1.0;
    t[1][0][1] = -1.0 / std::sqrt(2.0);
    // T(1,0) = [L-, T(1,1)] / sqrt(2), T(1,-1) = [L-, T(1,0)] / sqrt(2)
    auto lower = [](const cd (&a)[2][2], cd (&out)[2][2], double den) {
      // L- = [[0, 0], [1, 0]]: (L- a)[i][j] = (i == 1) ? a[0][j] : 0; (a L-)[i][j] = (j == 0) ? a[i][1] : 0
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          const cd la = (i == 1) ? a[0][j] : cd(0.0);
          const cd al = (j == 0) ? a[i][1] : cd(0.0);
          out[i][j] = (la - al) / den;
        }
    };
    lower(t[1], t[2], std::sqrt(2.0));
    lower(t[2], t[3], std::sqrt(2.0));
    double nrm[4];
    for (int k = 0; k < 4; k++) {
      double s = 0.0;
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) s += std::norm(t[k][i][j]);
      nrm[k] = std::sqrt(s);
    }
    auto mul = [](const cd (&a)[2][2], const cd (&b)[2][2], cd (&c)[2][2]) {
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) c[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j];
    };
    for (int n = 0; n < 4; n++)
      for (int m = 0; m < 4; m++) {
        cd nm[2][2], mn[2][2];
        mul(t[n], t[m], nm);
        mul(t[m], t[n], mn);
        for (int k = 0; k < 4; k++) {
          // norms[n] * tr((T_k/|T_k|)' (T_n/|T_n|)(T_m/|T_m|)), real part
          cd sl = 0.0, sr = 0.0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
              sl += std::conj(t[k][i][j]) * nm[i][j];
              sr += std::conj(t[k][i][j]) * mn[i][j];
            }
          const double f = 1.0 / (nrm[k] * nrm[m]);
          double l = sl.real() * f, r = sr.real() * f;
          left[n][m][k] = std::fabs(l) < 1e-14 ? 0.0 : l;
          right[n][m][k] = std::fabs(r) < 1e-14 ? 0.0 : r;
        }
      }
    for (int side = 0; side < 2; side++)
      for (int n = 0; n < 4; n++)
        for (int m = 0; m < 4; m++) {
          NzList& z = nz[side][n][m];
          for (int k = 0; k < 4; k++) {
            const double x = side == 0 ? left[n][m][k] : right[n][m][k];
            if (x != 0.0) {
              z.d[z.n] = k;
              z.c[z.n++] = x;
            }
          }
        }
  }
};

struct Term {
  int nspin;          // 1 or 2
  int spin[2];
  int ist[2];         // IST index of each active spin (1..3)
  double coef;        // product of the operator names' coefficients
  cd iso;             // isotropic coefficient
  cd istc[8], irr[8];  // rank-1 (0..2) and rank-2 (3..7) IST coefficients and irreducible components
};

// CSC of one matrix, values interleaved complex
struct Csc {
  int64_t n = 0;
  std::vector<uint64_t> jc, ir;
  std::vector<double> pr;
};

// I and the 34 Q slots (rank 1 (k,m) row-major, then rank 2), each a CSC matrix. Column by
// column (source state c): every term's left and right products from state c are scattered
// into dense per-slot row accumulators (coefficient times the term's weight in that slot: the
// isotropic coefficient for I, ist_k * irr_m for Q{r}{k,m}); the touched rows are then
// quantised to tol (clean_up.m: tol * round(x / tol), ties to even) and the nonzeros appended.
inline std::vector<Csc> build_hamiltonian(int n, const std::vector<int8_t>& digits, int64_t dim,
                                          const std::vector<Term>& terms, double tol) {
  if (n < 1 || n > 31) throw std::invalid_argument("the native builder takes 1 to 31 spins");
  std::vector<uint64_t> keys((size_t)dim);
  for (int64_t s = 0; s < dim; s++) {
    uint64_t k = 0;
    for (int a = 0; a < n; a++) {
      const int d = digits[(size_t)s * n + a];
      if (d < 0 || d > 3) throw std::invalid_argument("basis digits must be 0..3");
      k = k * 4 + (uint64_t)d;
    }
    keys[(size_t)s] = k;
  }
  for (int64_t s = 1; s < dim; s++)
    if (keys[(size_t)s] <= keys[(size_t)s - 1]) throw std::invalid_argument("basis rows must ascend, without repeats");
  static const IstTables tb;
  // the slots each term reaches, with its weight there
  struct Use {
    int slot;
    cd w;
  };
  std::vector<std::vector<Use>> uses(terms.size());
  for (size_t i = 0; i < terms.size(); i++) {
    const Term& t = terms[i];
    if (t.nspin < 1 || t.nspin > 2) throw std::invalid_argument("a term acts on one or two spins");
    for (int a = 0; a < t.nspin; a++)
      if (t.spin[a] < 0 || t.spin[a] >= n || t.ist[a] < 1 || t.ist[a] > 3)
        throw std::invalid_argument("term spin or IST index out of range");
    double sirr = 0.0, sist = 0.0;
    for (int k = 0; k < 8; k++) {
      sirr += std::abs(t.irr[k]);
      sist += std::abs(t.istc[k]);
    }
    if (std::abs(t.iso) < tol && (sirr < tol || sist < tol)) continue;  // hamiltonian.m: the term is dropped
    if (std::abs(t.iso) > tol) uses[i].push_back({0, t.iso});
    int slot = 1;
    for (int r = 1; r <= 2; r++) {
      const int off = (r == 1) ? 0 : 3;
      for (int k = 0; k < 2 * r + 1; k++)
        for (int m = 0; m < 2 * r + 1; m++, slot++) {
          const cd c = t.istc[off + k] * t.irr[off + m];
          if (std::abs(c) > tol) uses[i].push_back({slot, c});
        }
    }
  }
  const int ns = 35;
  std::vector<Csc> out((size_t)ns);
  std::vector<std::vector<cd>> acc((size_t)ns, std::vector<cd>((size_t)dim, cd(0.0)));
  std::vector<std::vector<char>> hit((size_t)ns, std::vector<char>((size_t)dim, 0));
  std::vector<std::vector<int64_t>> touched((size_t)ns);
  for (int s = 0; s < ns; s++) {
    out[(size_t)s].n = dim;
    out[(size_t)s].jc.assign((size_t)dim + 1, 0);
  }
  for (int64_t c = 0; c < dim; c++) {
    for (size_t i = 0; i < terms.size(); i++) {
      if (uses[i].empty()) continue;
      const Term& t = terms[i];
      int src[2] = {0, 0};
      uint64_t place[2] = {0, 0};
      for (int a = 0; a < t.nspin; a++) {
        src[a] = digits[(size_t)c * n + t.spin[a]];
        place[a] = (uint64_t)1 << (2 * (n - 1 - t.spin[a]));
      }
      for (int side = 0; side < 2; side++) {
        const double sgn = side == 0 ? 1.0 : -1.0;
        // the nonzero structure coefficients of each active spin from its source digit
        const NzList& l0 = tb.nz[side][t.ist[0]][src[0]];
        const NzList& l1 = tb.nz[side][t.nspin > 1 ? t.ist[1] : 0][t.nspin > 1 ? src[1] : 0];
        const int n1 = t.nspin > 1 ? l1.n : 1;
        for (int i0 = 0; i0 < l0.n; i0++)
          for (int i1 = 0; i1 < n1; i1++) {
            double coeff = l0.c[i0];
            uint64_t key = keys[(size_t)c] - (uint64_t)src[0] * place[0] + (uint64_t)l0.d[i0] * place[0];
            if (t.nspin > 1) {
              coeff *= l1.c[i1];
              key = key - (uint64_t)src[1] * place[1] + (uint64_t)l1.d[i1] * place[1];
            }
            auto it = std::lower_bound(keys.begin(), keys.end(), key);
            if (it == keys.end() || *it != key) continue;  // outside the basis
            const int64_t row = it - keys.begin();
            const double v = sgn * t.coef * coeff;
            for (const Use& u : uses[i]) {
              auto& h = hit[(size_t)u.slot];
              if (!h[(size_t)row]) {
                h[(size_t)row] = 1;
                touched[(size_t)u.slot].push_back(row);
              }
              acc[(size_t)u.slot][(size_t)row] += u.w * v;
            }
          }
      }
    }
    for (int s = 0; s < ns; s++) {
      auto& tr = touched[(size_t)s];
      std::sort(tr.begin(), tr.end());
      Csc& o = out[(size_t)s];
      for (int64_t row : tr) {
        const cd x = acc[(size_t)s][(size_t)row];
        acc[(size_t)s][(size_t)row] = 0.0;
        hit[(size_t)s][(size_t)row] = 0;
        const double re = tol * std::nearbyint(x.real() / tol), im = tol * std::nearbyint(x.imag() / tol);
        if (re == 0.0 && im == 0.0) continue;
        o.ir.push_back((uint64_t)row);
        o.pr.push_back(re);
        o.pr.push_back(im);
      }
      o.jc[(size_t)c + 1] = o.ir.size();
      tr.clear();
    }
  }
  return out;
}

}  // namespace spb

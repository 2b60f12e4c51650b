// The native Hamiltonian builder (spinach_amd/csrc/builder.hpp) built with AddressSanitizer and
// UndefinedBehaviorSanitizer. TEST INFRASTRUCTURE ONLY.
//   builder_san
// builds commutation superoperators whose structure the algebra fixes and checks them, then
// feeds the builder invalid inputs it must reject; one line per case, exit status 1 on a failed
// check.
//   ZEEMAN   [Lz(a), T] = m T on the full 4-spin basis: I is diagonal, the entry of a state is
//            the projection m of spin a's IST (digit 1 -> +1, 2 -> 0, 3 -> -1, 0 -> 0)
//   FLIPFLOP the commutation superoperator of Lx(0)Lx(1) + Ly(0)Ly(1), a Hermitian operator, on
//            the full basis is real symmetric and commutes with the total-Lz superoperator
//   PRUNED   the same on a basis with every fourth state removed: rows stay inside the basis
//   ANISO    a rank-2 term fills exactly the Q slots where ist_k * irr_m is nonzero
//   REJECT   invalid digits, unsorted rows, spin or IST index out of range, spin count
#include <cmath>
#include <cstdio>
#include <map>
#include <utility>
#include <vector>

#include "../../spinach_amd/csrc/builder.hpp"

namespace {

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    failures++;
  }
}

std::vector<int8_t> full_basis(int n) {
  const int64_t dim = (int64_t)1 << (2 * n);
  std::vector<int8_t> d((size_t)(dim * n));
  for (int64_t s = 0; s < dim; s++)
    for (int a = 0; a < n; a++) d[(size_t)(s * n + a)] = (int8_t)((s >> (2 * (n - 1 - a))) & 3);
  return d;
}

// CSC structure: column pointers ascend, rows inside [0, dim) and ascending within a column
bool valid(const spb::Csc& m) {
  if ((int64_t)m.jc.size() != m.n + 1 || m.jc[0] != 0 || m.jc.back() != m.ir.size() ||
      m.pr.size() != 2 * m.ir.size())
    return false;
  for (int64_t c = 0; c < m.n; c++) {
    if (m.jc[(size_t)c + 1] < m.jc[(size_t)c]) return false;
    for (uint64_t k = m.jc[(size_t)c]; k < m.jc[(size_t)c + 1]; k++) {
      if ((int64_t)m.ir[k] >= m.n) return false;
      if (k > m.jc[(size_t)c] && m.ir[k] <= m.ir[k - 1]) return false;
    }
  }
  return true;
}

std::map<std::pair<int64_t, int64_t>, spb::cd> entries(const spb::Csc& m) {
  std::map<std::pair<int64_t, int64_t>, spb::cd> e;
  for (int64_t c = 0; c < m.n; c++)
    for (uint64_t k = m.jc[(size_t)c]; k < m.jc[(size_t)c + 1]; k++)
      e[{(int64_t)m.ir[k], c}] = spb::cd(m.pr[2 * k], m.pr[2 * k + 1]);
  return e;
}

spb::Term term1(int spin, int ist, double coef, spb::cd iso) {
  spb::Term t{};
  t.nspin = 1;
  t.spin[0] = spin;
  t.ist[0] = ist;
  t.coef = coef;
  t.iso = iso;
  return t;
}

spb::Term term2(int s0, int i0, int s1, int i1, double coef, spb::cd iso) {
  spb::Term t = term1(s0, i0, coef, iso);
  t.nspin = 2;
  t.spin[1] = s1;
  t.ist[1] = i1;
  return t;
}

int projection(int digit) { return digit == 1 ? 1 : digit == 3 ? -1 : 0; }

void zeeman() {
  const int n = 4;
  std::vector<int8_t> d = full_basis(n);
  const int64_t dim = 256;
  for (int a = 0; a < n; a++) {
    auto out = spb::build_hamiltonian(n, d, dim, {term1(a, 2, 1.0, 1.0)}, 1e-10);
    check(out.size() == 35 && valid(out[0]), "ZEEMAN structure");
    auto e = entries(out[0]);
    int64_t nz = 0;
    for (int64_t s = 0; s < dim; s++) {
      const int m = projection(d[(size_t)(s * n + a)]);
      nz += m != 0;
      auto it = e.find({s, s});
      const spb::cd got = it == e.end() ? spb::cd(0.0) : it->second;
      check(std::abs(got - spb::cd((double)m)) < 1e-12, "ZEEMAN diagonal entry");
    }
    check((int64_t)e.size() == nz, "ZEEMAN off-diagonal entries");
    for (int q = 1; q < 35; q++) check(out[(size_t)q].ir.empty(), "ZEEMAN anisotropic slot not empty");
  }
  std::printf("ZEEMAN ok=%d\n", failures == 0);
}

// Lx Lx + Ly Ly = (L+ L- + L- L+) / 2; 'L+' = -sqrt(2) T(1,+1), 'L-' = sqrt(2) T(1,-1)
std::vector<spb::Term> flipflop(int a, int b) {
  return {term2(a, 1, b, 3, -2.0, 0.5), term2(a, 3, b, 1, -2.0, 0.5)};
}

void flipflop_case(const std::vector<int8_t>& d, int64_t dim, const char* tag) {
  const int n = 4;
  const int before = failures;
  auto ff = spb::build_hamiltonian(n, d, dim, flipflop(0, 1), 1e-10);
  check(valid(ff[0]), "FLIPFLOP structure");
  auto e = entries(ff[0]);
  for (auto& kv : e) {
    check(std::abs(kv.second.imag()) == 0.0, "FLIPFLOP entry not real");
    auto tr = e.find({kv.first.second, kv.first.first});
    check(tr != e.end() && std::abs(tr->second - kv.second) < 1e-12, "FLIPFLOP not symmetric");
  }
  // a flip-flop conserves the total projection: entries join states of equal total m
  for (auto& kv : e) {
    int mr = 0, mc = 0;
    for (int a = 0; a < n; a++) {
      mr += projection(d[(size_t)(kv.first.first * n + a)]);
      mc += projection(d[(size_t)(kv.first.second * n + a)]);
    }
    check(mr == mc, "FLIPFLOP joins states of different total projection");
  }
  check(!e.empty(), "FLIPFLOP empty");
  std::printf("%s dim=%lld nnz=%zu ok=%d\n", tag, (long long)dim, e.size(), failures == before);
}

void aniso() {
  const int n = 2;
  std::vector<int8_t> d = full_basis(n);
  const int before = failures;
  spb::Term t = term2(0, 2, 1, 2, 1.0, 0.0);
  t.istc[3 + 2] = 1.0;  // rank-2 IST component k = 2 (q = 0)
  t.irr[3 + 0] = spb::cd(0.5, 0.25);
  t.irr[3 + 4] = spb::cd(-0.5, 0.25);
  auto out = spb::build_hamiltonian(n, d, 16, {t}, 1e-10);
  // rank-2 slots follow the 9 rank-1 slots after I: slot = 1 + 9 + 5k + m
  for (int s = 0; s < 35; s++) {
    const bool want = (s == 1 + 9 + 5 * 2 + 0) || (s == 1 + 9 + 5 * 2 + 4);
    check(valid(out[(size_t)s]), "ANISO structure");
    check(out[(size_t)s].ir.empty() != want, "ANISO slot occupancy");
  }
  std::printf("ANISO ok=%d\n", failures == before);
}

void reject(const char* tag, int n, const std::vector<int8_t>& d, int64_t dim, const std::vector<spb::Term>& t) {
  try {
    spb::build_hamiltonian(n, d, dim, t, 1e-10);
    std::printf("FAIL %s accepted\n", tag);
    failures++;
  } catch (std::invalid_argument& e) {
    std::printf("REJECT %s: %s\n", tag, e.what());
  }
}

}  // namespace

int main() {
  zeeman();
  {
    std::vector<int8_t> d = full_basis(4);
    flipflop_case(d, 256, "FLIPFLOP");
    std::vector<int8_t> p;
    int64_t dim = 0;
    for (int64_t s = 0; s < 256; s++) {
      if (s % 4 == 3) continue;
      p.insert(p.end(), d.begin() + s * 4, d.begin() + s * 4 + 4);
      dim++;
    }
    flipflop_case(p, dim, "PRUNED");
  }
  aniso();
  std::vector<int8_t> d = full_basis(2);
  std::vector<spb::Term> ok = {term1(0, 2, 1.0, 1.0)};
  {
    std::vector<int8_t> bad = d;
    bad[5] = 4;
    reject("digit", 2, bad, 16, ok);
  }
  {
    std::vector<int8_t> bad = d;
    std::swap(bad[2], bad[4]);
    std::swap(bad[3], bad[5]);
    reject("order", 2, bad, 16, ok);
  }
  reject("spin", 2, d, 16, {term1(2, 2, 1.0, 1.0)});
  reject("ist", 2, d, 16, {term1(0, 4, 1.0, 1.0)});
  reject("nspin", 0, {}, 0, ok);
  std::printf("DONE failures=%d\n", failures);
  return failures ? 1 : 0;
}

// The host planner (spinach_amd/csrc/planner.hpp) built with AddressSanitizer
// and UndefinedBehaviorSanitizer. TEST INFRASTRUCTURE ONLY.
//   planner_san <problem.bin>
// plans the problem, then a set of corrupted copies of it that the planner
// must reject, and prints one line per case.
#include <cstdio>
#include <numeric>

#include "../../spinach_amd/csrc/planner.hpp"
#include "problem_io.hpp"

namespace {

int run(const Problem& p, std::vector<sp_csc>& m) {
  try {
    spp::SysPlan sp = spp::plan_system(&m[0], &m[1], p.al.data(), p.be.data(), p.ga.data(), p.w.data(),
                                       (int64_t)p.w.size(), p.rho0.data(), p.coil.data());
    std::vector<int32_t> start, col, pos;
    spp::SellDesc d{};
    spp::sell_layout(sp, start, col, pos, d);
    int32_t is_real = spp::generator_is_real_full(&m[0], &m[1]) ? 1 : 0;
    long long colsum = std::accumulate(col.begin(), col.end(), 0LL);
    std::printf("PLAN n=%d np=%d nnz=%d real=%d nslices=%d lanes=%d sell_nnz=%d colsum=%lld\n", sp.n, sp.np,
                sp.nnz, is_real, d.nslices, d.lanes, d.sell_nnz, colsum);
    return 0;
  } catch (spp::SpError& e) {
    std::printf("REJECT %d %s\n", e.code, e.what());
    return 1;
  }
}

std::vector<sp_csc> views(Problem& p) {
  std::vector<sp_csc> m(35);
  for (int t = 0; t < 35; t++) {
    m[t].n = p.mats[t].n ? p.mats[t].n : 0;
    m[t].jc = p.mats[t].jc.empty() ? nullptr : p.mats[t].jc.data();
    m[t].ir = p.mats[t].ir.empty() ? nullptr : p.mats[t].ir.data();
    m[t].pr = p.mats[t].pr.empty() ? nullptr : p.mats[t].pr.data();
    m[t].is_complex = 1;
  }
  return m;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  Problem p = read_problem(argv[1]);
  std::vector<sp_csc> m = views(p);
  run(p, m);
  // corrupted copies: each must be rejected with SP_EINVAL
  {
    Problem q = p;
    q.mats[0].ir[0] = (uint64_t)q.dim;  // row index out of range
    std::vector<sp_csc> v = views(q);
    run(q, v);
  }
  {
    Problem q = p;
    q.mats[0].jc[1] = q.mats[0].jc[q.dim] + 5;  // column pointers decrease
    std::vector<sp_csc> v = views(q);
    run(q, v);
  }
  {
    Problem q = p;
    q.rho0[0] = 0.0 / 0.0;  // non-finite state
    std::vector<sp_csc> v = views(q);
    run(q, v);
  }
  {
    Problem q = p;
    q.w.clear();
    q.al.clear();
    q.be.clear();
    q.ga.clear();  // empty grid
    std::vector<sp_csc> v = views(q);
    run(q, v);
  }
  {
    Problem q = p;
    q.mats[3].n = q.dim + 1;  // Q slot of the wrong dimension
    q.mats[3].jc.assign((size_t)q.dim + 2, 0);
    std::vector<sp_csc> v = views(q);
    run(q, v);
  }
  return 0;
}

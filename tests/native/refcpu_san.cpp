// The oracle (oracle/refcpu.cpp) built with AddressSanitizer and
// UndefinedBehaviorSanitizer. TEST INFRASTRUCTURE ONLY.
//   refcpu_san <problem.bin> <out.bin>
// runs refcpu_powder in its three variants, with the stepping branch forced,
// with a dead time and with path tracing forced (merge_dim 2), and writes the
// six FIDs (each nsteps+1 complex) to out.bin.
#include <cstdio>
#include <vector>

#include "problem_io.hpp"

extern "C" int refcpu_powder(int64_t dim, const int64_t* const* ptrs, const int64_t* const* idxs,
                             const double* const* vals, const double* alphas, const double* betas,
                             const double* gammas, const double* weights, int64_t n_or, const double* rho0,
                             const double* coil, double dt, int64_t nsteps, const double* tols, int variant,
                             int nthreads, double* fid_out, double* per_orient_out);
extern "C" const char* refcpu_last_error();

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  Problem p = read_problem(argv[1]);
  // the problem file holds CSC, the oracle takes CSR: convert each slot
  std::vector<std::vector<int64_t>> ptr(35), idx(35);
  std::vector<std::vector<double>> val(35);
  std::vector<const int64_t*> pp(35, nullptr), ii(35, nullptr);
  std::vector<const double*> vv(35, nullptr);
  for (int t = 0; t < 35; t++) {
    const ProblemMat& m = p.mats[t];
    if (m.n == 0) continue;
    const int64_t n = m.n;
    ptr[t].assign(n + 1, 0);
    for (uint64_t k = 0; k < m.jc[n]; k++) ptr[t][m.ir[k] + 1]++;
    for (int64_t r = 0; r < n; r++) ptr[t][r + 1] += ptr[t][r];
    idx[t].resize(m.jc[n]);
    val[t].resize(2 * m.jc[n]);
    std::vector<int64_t> next(ptr[t].begin(), ptr[t].end() - 1);
    for (int64_t c = 0; c < n; c++)
      for (uint64_t k = m.jc[c]; k < m.jc[c + 1]; k++) {
        const int64_t q = next[m.ir[k]]++;
        idx[t][q] = c;
        val[t][2 * q] = m.pr[2 * k];
        val[t][2 * q + 1] = m.pr[2 * k + 1];
      }
    pp[t] = ptr[t].data();
    ii[t] = idx[t].data();
    vv[t] = val[t].data();
  }
  const int64_t npts = p.nsteps + 1;
  const double base[11] = {1e-10, 1e-24, 1e-10, 200, 10000, 1e-10, 0.5, 32, 1000, 0, 0};
  struct Case {
    int variant;
    int tol_index;
    double tol_value;
  } cases[6] = {{0, -1, 0}, {1, -1, 0}, {2, -1, 0}, {1, 4, 2.0}, {0, 10, 3.3e-6}, {1, 8, 2.0}};
  std::vector<double> out;
  for (const Case& c : cases) {
    double tols[11];
    for (int k = 0; k < 11; k++) tols[k] = base[k];
    if (c.tol_index >= 0) tols[c.tol_index] = c.tol_value;
    std::vector<double> fid(2 * npts);
    if (refcpu_powder(p.dim, pp.data(), ii.data(), vv.data(), p.al.data(), p.be.data(), p.ga.data(), p.w.data(),
                      (int64_t)p.w.size(), p.rho0.data(), p.coil.data(), p.dt, p.nsteps, tols, c.variant, 1,
                      fid.data(), nullptr) != 0) {
      std::printf("FAIL %s\n", refcpu_last_error());
      return 1;
    }
    out.insert(out.end(), fid.begin(), fid.end());
  }
  FILE* f = std::fopen(argv[2], "wb");
  if (!f) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("OK\n");
  return 0;
}

// Reader of the problem files the sanitizer tests write (tests/test_boundary.py,
// write_problem). TEST INFRASTRUCTURE ONLY. Little endian:
//   u64 dim; 35 x matrix (slot 0 = I, then the 34 Q slots): u64 n (0 = empty),
//     then u64 nnz, (n+1) u64 jc, nnz u64 ir, 2*nnz doubles (re, im)  [CSC]
//   u64 n_or; n_or doubles each of alphas, betas, gammas, weights;
//   2*dim doubles rho0; 2*dim doubles coil; double dt; u64 nsteps
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

struct ProblemMat {
  int64_t n = 0;
  std::vector<uint64_t> jc, ir;
  std::vector<double> pr;
};

struct Problem {
  int64_t dim = 0;
  ProblemMat mats[35];
  std::vector<double> al, be, ga, w, rho0, coil;
  double dt = 0;
  int64_t nsteps = 0;
};

template <typename T>
inline void rd_vec(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short problem file");
}

inline uint64_t rd_u64(FILE* f) {
  uint64_t v;
  if (std::fread(&v, 8, 1, f) != 1) throw std::runtime_error("short problem file");
  return v;
}

inline Problem read_problem(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open problem file");
  Problem p;
  p.dim = (int64_t)rd_u64(f);
  for (auto& m : p.mats) {
    m.n = (int64_t)rd_u64(f);
    if (m.n == 0) continue;
    const uint64_t nnz = rd_u64(f);
    rd_vec(f, m.jc, (size_t)m.n + 1);
    rd_vec(f, m.ir, nnz);
    rd_vec(f, m.pr, 2 * nnz);
  }
  const uint64_t n_or = rd_u64(f);
  rd_vec(f, p.al, n_or);
  rd_vec(f, p.be, n_or);
  rd_vec(f, p.ga, n_or);
  rd_vec(f, p.w, n_or);
  rd_vec(f, p.rho0, 2 * (size_t)p.dim);
  rd_vec(f, p.coil, 2 * (size_t)p.dim);
  if (std::fread(&p.dt, 8, 1, f) != 1) throw std::runtime_error("short problem file");
  p.nsteps = (int64_t)rd_u64(f);
  std::fclose(f);
  return p;
}

// Recording stand-in for libspinach_powder's C-ABI (include/spinach_powder.h),
// linked into the MEX gateway's CPU test build. TEST INFRASTRUCTURE ONLY.
// It writes what the gateway passed (as JSON, to $SP_STUB_LOG) and returns
// known results: fid[t] = (t + 1) + 0.5i t and per[o][t] = o + 1e-3 t - 1i o,
// so that the test checks the gateway's marshalling in both directions.
// $SP_STUB_FAIL = create | plan | dead_time | execute | result makes that call
// fail with the message "stub failure in <call>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "spinach_powder.h"

struct sp_ctx {
  std::string err;
};
struct sp_plan {
  sp_ctx* ctx;
  int64_t n_or, nsteps;
  std::string log;
};

namespace {
std::string g_err = "stub: no error";

bool failing(const char* what) {
  const char* f = std::getenv("SP_STUB_FAIL");
  return f && std::strcmp(f, what) == 0;
}

void num(std::string& s, double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  s += b;
}

void arr(std::string& s, const double* v, size_t n) {
  s += "[";
  for (size_t i = 0; i < n; i++) {
    if (i) s += ",";
    num(s, v[i]);
  }
  s += "]";
}

// CSC as the test needs it: n, the nonzero count and every entry (row, col, re, im)
void csc(std::string& s, const sp_csc* m) {
  s += "{\"n\":";
  num(s, (double)m->n);
  s += ",\"entries\":[";
  if (m->n > 0 && m->jc) {
    bool first = true;
    for (int64_t c = 0; c < m->n; c++)
      for (uint64_t p = m->jc[c]; p < m->jc[c + 1]; p++) {
        if (!first) s += ",";
        first = false;
        s += "[";
        num(s, (double)m->ir[p]);
        s += ",";
        num(s, (double)c);
        s += ",";
        num(s, m->is_complex ? m->pr[2 * p] : m->pr[p]);
        s += ",";
        num(s, m->is_complex ? m->pr[2 * p + 1] : 0.0);
        s += "]";
      }
  }
  s += "]}";
}
}  // namespace

extern "C" {

sp_ctx* sp_create(int device) {
  if (failing("create") || device != 0) {
    g_err = "stub failure in create";
    return nullptr;
  }
  return new sp_ctx;
}

void sp_destroy(sp_ctx* ctx) { delete ctx; }

const char* sp_last_error(const sp_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

sp_plan* sp_plan_powder_relax(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas,
                              const double* betas, const double* gammas, const double* weights, int64_t n_or,
                              const double* rho0, const double* coil, double dt, int64_t nsteps,
                              const sp_tols* tols, int mode, const sp_relax* rlx) {
  sp_plan* p = sp_plan_powder(ctx, I, Q, alphas, betas, gammas, weights, n_or, rho0, coil, dt, nsteps, tols, mode);
  if (!p) return p;
  p->log += ",\"relax\":{\"R0\":";
  if (rlx && rlx->R0)
    csc(p->log, rlx->R0);
  else
    p->log += "null";
  p->log += ",\"diag\":";
  if (rlx && rlx->diag)
    arr(p->log, rlx->diag, 7 * (size_t)I->n);
  else
    p->log += "null";
  p->log += "}";
  return p;
}

sp_plan* sp_plan_powder(sp_ctx* ctx, const sp_csc* I, const sp_csc* Q, const double* alphas, const double* betas,
                        const double* gammas, const double* weights, int64_t n_or, const double* rho0,
                        const double* coil, double dt, int64_t nsteps, const sp_tols* tols, int mode) {
  if (failing("plan")) {
    ctx->err = "stub failure in plan";
    return nullptr;
  }
  auto* p = new sp_plan{ctx, n_or, nsteps, ""};
  std::string& s = p->log;
  s = "{\"I\":";
  csc(s, I);
  s += ",\"Q\":[";
  for (int q = 0; q < SP_NQ; q++) {
    if (q) s += ",";
    csc(s, &Q[q]);
  }
  s += "],\"alphas\":";
  arr(s, alphas, (size_t)n_or);
  s += ",\"betas\":";
  arr(s, betas, (size_t)n_or);
  s += ",\"gammas\":";
  arr(s, gammas, (size_t)n_or);
  s += ",\"weights\":";
  arr(s, weights, (size_t)n_or);
  s += ",\"rho0\":";
  arr(s, rho0, 2 * (size_t)I->n);
  s += ",\"coil\":";
  arr(s, coil, 2 * (size_t)I->n);
  s += ",\"dt\":";
  num(s, dt);
  s += ",\"nsteps\":";
  num(s, (double)nsteps);
  const double t[5] = {tols->liouv_zero, tols->zte_tol, tols->prop_chop, tols->small_matrix, tols->krylov_tol};
  s += ",\"tols\":";
  arr(s, t, 5);
  s += ",\"mode\":";
  num(s, mode);
  return p;
}

int sp_plan_set_dead_time(sp_plan* plan, double dead_time) {
  if (failing("dead_time")) {
    plan->ctx->err = "stub failure in dead_time";
    return SP_EINVAL;
  }
  plan->log += ",\"dead_time\":";
  num(plan->log, dead_time);
  return SP_OK;
}

int sp_execute(sp_plan* plan, sp_stats* stats) {
  if (failing("execute")) {
    plan->ctx->err = "stub failure in execute";
    return SP_ERANGE;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return SP_OK;
}

int sp_plan_result(sp_plan* plan, double* fid_out, double* per_orient_out) {
  if (failing("result")) {
    plan->ctx->err = "stub failure in result";
    return SP_EDEVICE;
  }
  const int64_t npts = plan->nsteps + 1;
  if (fid_out)
    for (int64_t t = 0; t < npts; t++) {
      fid_out[2 * t] = (double)(t + 1);
      fid_out[2 * t + 1] = 0.5 * (double)t;
    }
  if (per_orient_out)
    for (int64_t o = 0; o < plan->n_or; o++)
      for (int64_t t = 0; t < npts; t++) {
        per_orient_out[2 * (o * npts + t)] = (double)o + 1e-3 * (double)t;
        per_orient_out[2 * (o * npts + t) + 1] = -(double)o;
      }
  plan->log += ",\"fid_out\":";
  num(plan->log, fid_out ? 1 : 0);
  plan->log += ",\"per_out\":";
  num(plan->log, per_orient_out ? 1 : 0);
  return SP_OK;
}

void sp_plan_destroy(sp_plan* plan) {
  if (!plan) return;
  if (const char* path = std::getenv("SP_STUB_LOG")) {
    if (FILE* f = std::fopen(path, "w")) {
      std::fprintf(f, "%s}\n", plan->log.c_str());
      std::fclose(f);
    }
  }
  delete plan;
}

}  // extern "C"

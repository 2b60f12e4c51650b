// powder_hip.cpp - MEX gateway from Spinach's powder context to
// libspinach_powder (the GPU orientation loop and weighted sum). This is the
// file a maintainer adds as etc/mex/powder_hip.cpp; it is compiled like the
// other Spinach MEX files (etc/mex/compile_mex.m:16-33):
//   mex('-R2018a','-O','-DNDEBUG','-I<repo>/include','powder_hip.cpp',
//       '-L<repo>/spinach_amd','-lspinach_powder')
//
//   [fid,per]=powder_hip(I,Q,grid,rho0,coil,dt,nsteps,tols,opts)
//
//   I      isotropic Liouvillian, sparse (hamiltonian.m:718-738 after frqoffset)
//   Q      {Q1,Q2}: 3x3 and 5x5 cells of sparse matrices (hamiltonian.m:741-792)
//   grid   struct with alphas, betas, gammas, weights (powder.m:173-179)
//   rho0, coil  state vectors, full or sparse (state.m returns sparse)
//   dt, nsteps  1/parameters.sweep and parameters.npoints-1 (acquire.m:70-71)
//   tols   spin_system.tols (liouv_zero, zte_tol, prop_chop, small_matrix, krylov_tol)
//   opts   struct, every field optional:
//            precision  'double' (default) or 'single'
//            dead_time  parameters.dead_time (acquire.m:65-67)
//            sum_up     parameters.sum_up (powder.m:280-294), default true
//            krylov     true when spin_system.sys.enable contains 'krylov'
//            device     GPU index, default 0
//            R0         sparse relaxation superoperator, orientation-independent
//                       part (relaxation.m at any orientation, minus rlx_diag)
//            rlx_diag   dim x 7 real matrix [g0 g1 ... g6]: the diagonal of
//                       R(Omega) is g0 + sum_p q_p(Omega) g_p with
//                       q = (ox^2, oy^2, oz^2, 2ox*oy, 2ox*oz, 2oy*oz) of
//                       ort = [0 0 1]*euler2dcm(alpha,beta,gamma)
//                       (rlx_t1_t2.m rate tensors, relaxation.m 'damp');
//                       with R0 or rlx_diag the generator is L = H + 1i*R
//                       (evolution.m) on every path, the dead time included
//   fid    npoints x 1 complex, the weighted sum (sum_up true)
//   per    npoints x n_or complex, unweighted per-orientation FIDs (requested
//          as a second output, or the first output when sum_up is false;
//          powder.m turns its columns into the answer cell)
//
// Errors follow etc/mex/prune_cpu.cpp:199-223 ("Spinach:powder_hip:<id>").
// mexErrMsgIdAndTxt does not return (it unwinds with a longjmp), so it is
// called only from mexFunction itself, after run_gateway has returned and every
// C++ object of the call has been destroyed.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"
#include "spinach_powder.h"

namespace {

// One context per device for the life of the MATLAB session (mexAtExit).
sp_ctx* g_ctx[64] = {};
void release_contexts() {
  for (sp_ctx*& c : g_ctx) {
    sp_destroy(c);
    c = nullptr;
  }
}

struct GatewayError {  // plain data: no destructor to skip
  char id[64];
  char msg[1024];
};

struct Fail {  // thrown inside run_gateway, turned into GatewayError
  std::string id, msg;
};

[[noreturn]] void fail(const char* id, const std::string& msg) { throw Fail{id, msg}; }

bool is_sparse_double(const mxArray* m) { return m && mxIsSparse(m) && mxIsDouble(m); }

sp_csc as_csc(const mxArray* m, const char* what) {
  sp_csc s{};
  if (m == nullptr || mxIsEmpty(m)) return s;
  if (!is_sparse_double(m)) fail("input", std::string(what) + " must be a sparse double matrix.");
  if (mxGetM(m) != mxGetN(m)) fail("input", std::string(what) + " must be square.");
  s.n = (int64_t)mxGetN(m);
  s.jc = reinterpret_cast<const uint64_t*>(mxGetJc(m));
  s.ir = reinterpret_cast<const uint64_t*>(mxGetIr(m));
  s.is_complex = mxIsComplex(m) ? 1 : 0;
  s.pr = s.is_complex ? reinterpret_cast<const double*>(mxGetComplexDoubles(m)) : mxGetDoubles(m);
  return s;
}

// Column vector of length n, full or sparse, real or complex -> interleaved complex.
std::vector<double> as_cvec(const mxArray* v, size_t n, const char* what) {
  if (!v || !mxIsDouble(v)) fail("input", std::string(what) + " must be a double vector.");
  if (mxGetNumberOfElements(v) != n || mxGetN(v) != 1)
    fail("input", std::string(what) + " must be a column vector of the Liouville space dimension.");
  std::vector<double> out(2 * n, 0.0);
  if (mxIsSparse(v)) {
    const mwIndex* jc = mxGetJc(v);
    const mwIndex* ir = mxGetIr(v);
    const bool cplx = mxIsComplex(v);
    const mxComplexDouble* c = cplx ? mxGetComplexDoubles(v) : nullptr;
    const double* r = cplx ? nullptr : mxGetDoubles(v);
    for (mwIndex p = jc[0]; p < jc[1]; p++) {
      if (ir[p] >= n) fail("input", std::string(what) + ": sparse row index out of range.");
      out[2 * ir[p]] = cplx ? c[p].real : r[p];
      out[2 * ir[p] + 1] = cplx ? c[p].imag : 0.0;
    }
  } else if (mxIsComplex(v)) {
    const mxComplexDouble* c = mxGetComplexDoubles(v);
    for (size_t i = 0; i < n; i++) {
      out[2 * i] = c[i].real;
      out[2 * i + 1] = c[i].imag;
    }
  } else {
    const double* r = mxGetDoubles(v);
    for (size_t i = 0; i < n; i++) out[2 * i] = r[i];
  }
  return out;
}

double scalar(const mxArray* a, const char* what) {
  if (!a || !mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) != 1)
    fail("input", std::string(what) + " must be a real double scalar.");
  return mxGetScalar(a);
}

const mxArray* field(const mxArray* s, const char* name) {
  return (s && mxIsStruct(s)) ? mxGetField(s, 0, name) : nullptr;
}

std::vector<double> grid_column(const mxArray* grid, const char* name, size_t& n) {
  const mxArray* f = field(grid, name);
  if (!f || !mxIsDouble(f) || mxIsComplex(f) || mxIsSparse(f))
    fail("input", std::string("grid.") + name + " must be a real double vector.");
  const size_t k = mxGetNumberOfElements(f);
  if (n == 0) n = k;
  if (k != n || k == 0) fail("input", "grid.alphas, betas, gammas and weights must have the same positive length.");
  const double* d = mxGetDoubles(f);
  return std::vector<double>(d, d + k);
}

bool flag(const mxArray* opts, const char* name, bool dflt) {
  const mxArray* f = field(opts, name);
  if (!f) return dflt;
  if (mxGetNumberOfElements(f) != 1 || !(mxIsLogical(f) || mxIsDouble(f)))
    fail("input", std::string("opts.") + name + " must be a logical scalar.");
  return mxGetScalar(f) != 0.0;
}

sp_tols read_tols(const mxArray* t) {
  // tolerances.m defaults for fields the struct does not carry
  sp_tols out{1e-10, 1e-24, 1e-10, 200, 10000};
  if (!t || mxIsEmpty(t)) return out;
  if (!mxIsStruct(t)) fail("input", "tols must be the spin_system.tols structure.");
  struct {
    const char* name;
    double* dst;
  } f[] = {{"liouv_zero", &out.liouv_zero}, {"zte_tol", &out.zte_tol}, {"prop_chop", &out.prop_chop},
           {"small_matrix", &out.small_matrix}, {"krylov_tol", &out.krylov_tol}};
  for (auto& e : f)
    if (const mxArray* v = field(t, e.name)) *e.dst = scalar(v, (std::string("tols.") + e.name).c_str());
  return out;
}

// Owns a plan for the duration of the call (destroyed before any error is raised).
struct PlanGuard {
  sp_plan* p = nullptr;
  ~PlanGuard() { sp_plan_destroy(p); }
};

// The whole call; returns false with err filled on failure. Outputs are
// assigned to plhs only on success.
bool run_gateway(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[], GatewayError& err) {
  mxArray* out0 = nullptr;
  mxArray* out1 = nullptr;
  try {
    if (nrhs != 9) fail("nrhs", "nine inputs required: I, Q, grid, rho0, coil, dt, nsteps, tols, opts.");
    if (nlhs > 2) fail("nlhs", "at most two outputs: [fid,per].");
    const sp_csc I = as_csc(prhs[0], "I");
    if (I.n <= 0) fail("input", "I must be a non-empty sparse matrix.");
    const mxArray* Qc = prhs[1];
    if (!Qc || !mxIsCell(Qc) || mxGetNumberOfElements(Qc) != 2) fail("input", "Q must be a 1x2 cell {Q1,Q2}.");
    const mxArray* q1 = mxGetCell(Qc, 0);
    const mxArray* q2 = mxGetCell(Qc, 1);
    if (!q1 || !q2 || !mxIsCell(q1) || !mxIsCell(q2) || mxGetM(q1) != 3 || mxGetN(q1) != 3 || mxGetM(q2) != 5 ||
        mxGetN(q2) != 5)
      fail("input", "Q{1} must be a 3x3 cell and Q{2} a 5x5 cell.");
    sp_csc Q[SP_NQ] = {};
    for (int k = 0; k < 3; k++)  // slot order (k,m) row-major; MATLAB cells are column-major
      for (int m = 0; m < 3; m++) Q[k * 3 + m] = as_csc(mxGetCell(q1, k + 3 * m), "Q{1}{k,m}");
    for (int k = 0; k < 5; k++)
      for (int m = 0; m < 5; m++) Q[9 + k * 5 + m] = as_csc(mxGetCell(q2, k + 5 * m), "Q{2}{k,m}");
    size_t n_or = 0;
    const mxArray* grid = prhs[2];
    if (!grid || !mxIsStruct(grid)) fail("input", "grid must be a structure.");
    std::vector<double> al = grid_column(grid, "alphas", n_or), be = grid_column(grid, "betas", n_or),
                        ga = grid_column(grid, "gammas", n_or), w = grid_column(grid, "weights", n_or);
    const size_t dim = (size_t)I.n;
    std::vector<double> rho0 = as_cvec(prhs[3], dim, "rho0"), coil = as_cvec(prhs[4], dim, "coil");
    const double dt = scalar(prhs[5], "dt");
    const double ns = scalar(prhs[6], "nsteps");
    if (!(ns >= 0) || ns != (double)(int64_t)ns) fail("input", "nsteps must be a non-negative integer.");
    const int64_t nsteps = (int64_t)ns;
    const sp_tols tols = read_tols(prhs[7]);
    const mxArray* opts = prhs[8];
    if (opts && !mxIsEmpty(opts) && !mxIsStruct(opts)) fail("input", "opts must be a structure.");
    int mode = SP_MODE_FP64;
    if (const mxArray* pr = field(opts, "precision")) {
      char buf[16] = {0};
      if (!mxIsChar(pr) || mxGetString(pr, buf, sizeof(buf)) != 0) fail("input", "opts.precision must be a string.");
      if (std::strcmp(buf, "single") == 0)
        mode = SP_MODE_FP32;
      else if (std::strcmp(buf, "double") != 0)
        fail("input", "opts.precision must be 'double' or 'single'.");
    }
    if (flag(opts, "krylov", false)) mode |= SP_MODE_KRYLOV;
    const bool sum_up = flag(opts, "sum_up", true);
    double dead_time = 0.0;
    if (const mxArray* d = field(opts, "dead_time")) dead_time = scalar(d, "opts.dead_time");
    int device = 0;
    if (const mxArray* d = field(opts, "device")) device = (int)scalar(d, "opts.device");
    if (device < 0 || device >= 64) fail("device", "opts.device is out of range.");
    sp_csc R0{};
    if (const mxArray* r = field(opts, "R0")) {
      R0 = as_csc(r, "opts.R0");
      if (R0.n != 0 && R0.n != I.n) fail("input", "opts.R0 must be of the Liouville space dimension.");
    }
    const double* rlx_diag = nullptr;
    if (const mxArray* d = field(opts, "rlx_diag")) {
      if (!mxIsDouble(d) || mxIsComplex(d) || mxIsSparse(d) || mxGetM(d) != dim || mxGetN(d) != 7)
        fail("input", "opts.rlx_diag must be a real full dim x 7 matrix [g0 g1 ... g6].");
      rlx_diag = mxGetDoubles(d);  // column p = g_p, the library's p*dim layout
    }
    const sp_relax rlx{R0.n > 0 ? &R0 : nullptr, rlx_diag};
    const bool relax = rlx.R0 || rlx.diag;

    if (!g_ctx[device]) {
      g_ctx[device] = sp_create(device);
      if (!g_ctx[device]) fail("device", std::string("no GPU context: ") + sp_last_error(nullptr));
      mexAtExit(release_contexts);
    }
    sp_ctx* ctx = g_ctx[device];
    PlanGuard plan;
    plan.p = relax ? sp_plan_powder_relax(ctx, &I, Q, al.data(), be.data(), ga.data(), w.data(), (int64_t)n_or,
                                          rho0.data(), coil.data(), dt, nsteps, &tols, mode, &rlx)
                   : sp_plan_powder(ctx, &I, Q, al.data(), be.data(), ga.data(), w.data(), (int64_t)n_or,
                                    rho0.data(), coil.data(), dt, nsteps, &tols, mode);
    if (!plan.p) fail("plan", sp_last_error(ctx));
    if (dead_time != 0.0 && sp_plan_set_dead_time(plan.p, dead_time) != SP_OK) fail("input", sp_last_error(ctx));
    if (sp_execute(plan.p, nullptr) != SP_OK) fail("run", sp_last_error(ctx));
    const bool want_per = !sum_up || nlhs > 1;
    const mwSize npts = (mwSize)(nsteps + 1);
    if (sum_up) out0 = mxCreateDoubleMatrix(npts, 1, mxCOMPLEX);
    if (want_per) {
      mxArray* per = mxCreateDoubleMatrix(npts, (mwSize)n_or, mxCOMPLEX);
      if (sum_up)
        out1 = per;
      else
        out0 = per;
    }
    if ((sum_up && !out0) || (want_per && !(sum_up ? out1 : out0))) fail("alloc", "failed to allocate the outputs.");
    double* fid = sum_up ? reinterpret_cast<double*>(mxGetComplexDoubles(out0)) : nullptr;
    // orientation-major FIDs = columns of the npoints x n_or MATLAB matrix
    double* per = want_per ? reinterpret_cast<double*>(mxGetComplexDoubles(sum_up ? out1 : out0)) : nullptr;
    std::vector<double> scratch_fid;
    if (!fid) {
      scratch_fid.assign(2 * (size_t)npts, 0.0);
      fid = scratch_fid.data();
    }
    if (sp_plan_result(plan.p, fid, per) != SP_OK) fail("run", sp_last_error(ctx));
    plhs[0] = out0;
    if (nlhs > 1) plhs[1] = out1;
    else if (out1) mxDestroyArray(out1);
    return true;
  } catch (const Fail& f) {
    std::snprintf(err.id, sizeof(err.id), "Spinach:powder_hip:%s", f.id.c_str());
    std::snprintf(err.msg, sizeof(err.msg), "%s", f.msg.c_str());
  } catch (const std::bad_alloc&) {
    std::snprintf(err.id, sizeof(err.id), "Spinach:powder_hip:badAlloc");
    std::snprintf(err.msg, sizeof(err.msg), "Memory allocation failed.");
  } catch (const std::exception& e) {
    std::snprintf(err.id, sizeof(err.id), "Spinach:powder_hip:cppException");
    std::snprintf(err.msg, sizeof(err.msg), "C++ exception: %s", e.what());
  }
  if (out0) mxDestroyArray(out0);
  if (out1) mxDestroyArray(out1);
  return false;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  GatewayError err;
  if (!run_gateway(nlhs, plhs, nrhs, prhs, err)) mexErrMsgIdAndTxt(err.id, "%s", err.msg);
}

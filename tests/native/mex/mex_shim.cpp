// Implementation of the mex.h stand-in plus a driver that runs one gateway
// call. TEST INFRASTRUCTURE ONLY.
//
//   mex_driver <case.bin> <out.bin> <nlhs>
// reads the right-hand sides from case.bin, calls mexFunction, writes the
// left-hand sides to out.bin and prints "OK", or prints "ERR <id> <message>"
// when the gateway raised an error (mexErrMsgIdAndTxt longjmps back here, as
// MATLAB unwinds a MEX call). Before exiting it runs the mexAtExit handlers
// and destroys every array, so a leak or a double free shows under
// AddressSanitizer.
//
// Serialised array (little endian): u8 kind (0 dense double, 1 sparse double,
// 2 cell, 3 struct, 4 char, 5 logical), u64 m, u64 n, u8 complex, then
//   dense: m*n doubles, interleaved (re, im) when complex
//   sparse: u64 nnz, (n+1) u64 jc, nnz u64 ir, nnz values (interleaved if complex)
//   cell: m*n arrays, column-major
//   struct (1x1): u64 nfields, per field u64 name length, name bytes, array
//   char: m*n bytes; logical: m*n bytes
// A file is u64 count followed by that many arrays.
#include <csetjmp>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "mex.h"

struct mxArray_tag {
  int kind = 0;
  size_t m = 0, n = 0;
  bool cplx = false;
  std::vector<double> data;  // dense or sparse values, interleaved when complex
  std::vector<mwIndex> jc, ir;
  std::vector<mxArray*> cells;
  std::vector<std::pair<std::string, mxArray*>> fields;
  std::string text;
  std::vector<unsigned char> logic;
  ~mxArray_tag() {
    for (auto* c : cells) delete c;
    for (auto& f : fields) delete f.second;
  }
};

namespace {
std::jmp_buf g_jmp;
std::string g_err_id, g_err_msg;
std::vector<void (*)(void)> g_atexit;
}  // namespace

extern "C" {
bool mxIsSparse(const mxArray* a) { return a && a->kind == 1; }
bool mxIsDouble(const mxArray* a) { return a && (a->kind == 0 || a->kind == 1); }
bool mxIsComplex(const mxArray* a) { return a && a->cplx; }
bool mxIsEmpty(const mxArray* a) { return !a || a->m == 0 || a->n == 0; }
bool mxIsCell(const mxArray* a) { return a && a->kind == 2; }
bool mxIsStruct(const mxArray* a) { return a && a->kind == 3; }
bool mxIsChar(const mxArray* a) { return a && a->kind == 4; }
bool mxIsLogical(const mxArray* a) { return a && a->kind == 5; }
size_t mxGetM(const mxArray* a) { return a ? a->m : 0; }
size_t mxGetN(const mxArray* a) { return a ? a->n : 0; }
size_t mxGetNumberOfElements(const mxArray* a) { return a ? a->m * a->n : 0; }
mwIndex* mxGetJc(const mxArray* a) { return (a && a->kind == 1) ? const_cast<mwIndex*>(a->jc.data()) : nullptr; }
mwIndex* mxGetIr(const mxArray* a) { return (a && a->kind == 1) ? const_cast<mwIndex*>(a->ir.data()) : nullptr; }
double* mxGetDoubles(const mxArray* a) {
  return (a && mxIsDouble(a) && !a->cplx) ? const_cast<double*>(a->data.data()) : nullptr;
}
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
  return (a && mxIsDouble(a) && a->cplx) ? reinterpret_cast<mxComplexDouble*>(const_cast<double*>(a->data.data()))
                                         : nullptr;
}
double mxGetScalar(const mxArray* a) {
  if (!a) return 0.0;
  if (a->kind == 5) return a->logic.empty() ? 0.0 : (double)a->logic[0];
  if (a->kind == 1) return (a->ir.empty() || a->ir[0] != 0) ? 0.0 : a->data[0];
  return a->data.empty() ? 0.0 : a->data[0];
}
mxArray* mxGetCell(const mxArray* a, mwIndex i) { return (a && a->kind == 2 && i < a->cells.size()) ? a->cells[i] : nullptr; }
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name) {
  if (!a || a->kind != 3 || i != 0) return nullptr;
  for (auto& f : a->fields)
    if (f.first == name) return f.second;
  return nullptr;
}
int mxGetString(const mxArray* a, char* buf, mwSize len) {
  if (!a || a->kind != 4 || len == 0) return 1;
  const size_t k = std::min(a->text.size(), (size_t)len - 1);
  std::memcpy(buf, a->text.data(), k);
  buf[k] = 0;
  return a->text.size() < len ? 0 : 1;
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  auto* a = new mxArray_tag;
  a->kind = 0;
  a->m = m;
  a->n = n;
  a->cplx = (c == mxCOMPLEX);
  a->data.assign((a->cplx ? 2 : 1) * m * n, 0.0);
  return a;
}
void mxDestroyArray(mxArray* a) { delete a; }
int mexAtExit(void (*fn)(void)) {
  g_atexit.push_back(fn);
  return 0;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err_id = id;
  g_err_msg = buf;
  std::longjmp(g_jmp, 1);
}
}

namespace {

template <typename T>
bool rd(FILE* f, T& v) { return std::fread(&v, sizeof(T), 1, f) == 1; }

mxArray* read_array(FILE* f) {
  uint8_t kind, cplx;
  uint64_t m, n;
  if (!rd(f, kind) || !rd(f, m) || !rd(f, n) || !rd(f, cplx)) return nullptr;
  std::unique_ptr<mxArray_tag> a(new mxArray_tag);
  a->kind = kind;
  a->m = m;
  a->n = n;
  a->cplx = cplx != 0;
  const size_t w = a->cplx ? 2 : 1;
  if (kind == 0) {
    a->data.resize(w * m * n);
    if (std::fread(a->data.data(), sizeof(double), a->data.size(), f) != a->data.size()) return nullptr;
  } else if (kind == 1) {
    uint64_t nnz;
    if (!rd(f, nnz)) return nullptr;
    a->jc.resize(n + 1);
    a->ir.resize(nnz);
    a->data.resize(w * nnz);
    if (std::fread(a->jc.data(), sizeof(mwIndex), n + 1, f) != n + 1) return nullptr;
    if (std::fread(a->ir.data(), sizeof(mwIndex), nnz, f) != nnz) return nullptr;
    if (std::fread(a->data.data(), sizeof(double), a->data.size(), f) != a->data.size()) return nullptr;
  } else if (kind == 2) {
    for (size_t k = 0; k < m * n; k++) {
      mxArray* c = read_array(f);
      if (!c) return nullptr;
      a->cells.push_back(c);
    }
  } else if (kind == 3) {
    uint64_t nf;
    if (!rd(f, nf)) return nullptr;
    for (uint64_t k = 0; k < nf; k++) {
      uint64_t len;
      if (!rd(f, len)) return nullptr;
      std::string name(len, '\0');
      if (std::fread(&name[0], 1, len, f) != len) return nullptr;
      mxArray* c = read_array(f);
      if (!c) return nullptr;
      a->fields.emplace_back(name, c);
    }
  } else if (kind == 4) {
    a->text.resize(m * n);
    if (m * n > 0 && std::fread(&a->text[0], 1, m * n, f) != m * n) return nullptr;
  } else if (kind == 5) {
    a->logic.resize(m * n);
    if (m * n > 0 && std::fread(a->logic.data(), 1, m * n, f) != m * n) return nullptr;
  } else {
    return nullptr;
  }
  return a.release();
}

void write_array(FILE* f, const mxArray* a) {
  const uint8_t kind = 0, cplx = a->cplx ? 1 : 0;  // the gateways return dense doubles only
  const uint64_t m = a->m, n = a->n;
  std::fwrite(&kind, 1, 1, f);
  std::fwrite(&m, 8, 1, f);
  std::fwrite(&n, 8, 1, f);
  std::fwrite(&cplx, 1, 1, f);
  std::fwrite(a->data.data(), sizeof(double), a->data.size(), f);
}

// The gateway call, kept out of main's frame: a longjmp from mexErrMsgIdAndTxt
// lands in this function's setjmp.
int call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs) {
  if (setjmp(g_jmp) != 0) return 1;
  mexFunction(nlhs, plhs, nrhs, prhs);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: mex_driver case.bin out.bin nlhs\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  uint64_t count = 0;
  if (!rd(f, count) || count > 64) return 2;
  std::vector<mxArray*> rhs;
  for (uint64_t k = 0; k < count; k++) {
    mxArray* a = read_array(f);
    if (!a) {
      std::fprintf(stderr, "bad case file\n");
      return 2;
    }
    rhs.push_back(a);
  }
  std::fclose(f);
  const int nlhs = std::atoi(argv[3]);
  std::vector<mxArray*> lhs(std::max(1, nlhs), nullptr);
  const int rc = call(nlhs, lhs.data(), (int)rhs.size(), const_cast<const mxArray**>(rhs.data()));
  if (rc == 0) {
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    uint64_t nout = 0;
    for (int k = 0; k < nlhs; k++)
      if (lhs[k]) nout++;
    std::fwrite(&nout, 8, 1, o);
    for (int k = 0; k < nlhs; k++)
      if (lhs[k]) write_array(o, lhs[k]);
    std::fclose(o);
    std::printf("OK\n");
  } else {
    std::printf("ERR %s %s\n", g_err_id.c_str(), g_err_msg.c_str());
  }
  for (auto* a : lhs) delete a;
  for (auto* a : rhs) delete a;
  for (auto fn = g_atexit.rbegin(); fn != g_atexit.rend(); ++fn) (*fn)();
  std::fflush(stdout);
  return 0;
}

// Host build of the divide-and-conquer eigensolver of the spectral path: the
// same phase functions as the k_dceig kernel, run phase by phase with a loop
// over thread ids (a phase's threads touch disjoint data, so this order is
// one valid schedule). Built by tests/test_dc_host.py; not part of the library.
#include "../../spinach_amd/csrc/spectral_kernels.hpp"

using namespace spk;

template <int NM>
static int dc_host_solve_t(const double* in, int n, double* out, int* status, int* levels) {
  static DcSharedT<NM> S;
  const int T = DcSharedT<NM>::kThr;
  for (int i = 0; i < 6 * NM; i++) out[i] = in[i];
  for (int t = 0; t < T; t++) dc_phase_load(S, out, NM, t, T);
  dc_phase_plan(S, n);
  *levels = S.nlev;
  for (int t = 0; t < T; t++) dc_phase_leaf(S, t);
  for (int l = S.nlev - 1; l >= 0; l--) {
    for (int t = 0; t < T; t++) dc_phase_m1(S, l, n, t);
    for (int t = 0; t < T; t++) dc_phase_m2(S, l, t);
    for (int t = 0; t < T; t++) dc_phase_m3a(S, l, n, t);
    for (int t = 0; t < T; t++) dc_phase_m3b(S, l, t);
    for (int t = 0; t < T; t++) dc_phase_m4(S, l, n, t);
    for (int t = 0; t < T; t++) dc_phase_m5(S, l, n, t);
    if (l > 0) {
      for (int t = 0; t < T; t++) dc_phase_m6a(S, l, n, t);
      for (int t = 0; t < T; t++) dc_phase_m6b(S, t);
    }
  }
  for (int t = 0; t < T; t++) dc_phase_store(S, out, NM, n, t, T);
  int mx = 0;
  for (int i = 0; i < NM; i++) mx = S.its[i] > mx ? S.its[i] : mx;
  *status = S.fail ? -1 : mx;
  return 0;
}

// in/out: 6 * nm doubles (d | e | four vectors), nm = kTriNP (n <= kTriNP, the solver sized for
// the 256-thread tridiagonalisation), kTriWNP or kTriGNP (the ones sized for k_tridiag_w)
extern "C" int dc_host_solve(const double* in, int n, int nm, double* out, int* status, int* levels) {
  if (nm == kTriNP && n >= 0 && n <= kTriNP) return dc_host_solve_t<kTriNP>(in, n, out, status, levels);
  if (nm == kTriWNP && n >= 0 && n <= kTriWNP) return dc_host_solve_t<kTriWNP>(in, n, out, status, levels);
  if (nm == kTriGNP && n >= 0 && n <= kTriGNP) return dc_host_solve_t<kTriGNP>(in, n, out, status, levels);
  return 1;
}

// the per-slot Wigner coefficients (one thread per slot on the device) against
// the whole-set function, bit for bit: returns the number of differing entries
extern "C" int wigner_slot_mismatches(double a, double b, double g) {
  double cre[34], cim[34], sre[34], sim[34];
  spk::wigner_coeffs(a, b, g, cre, cim);
  for (int t = 0; t < 34; t++) spk::wigner_coeff_slot(a, b, g, t, sre, sim);
  int bad = 0;
  for (int t = 0; t < 34; t++) bad += (cre[t] != sre[t]) + (cim[t] != sim[t]);
  return bad;
}

// dc_select_bit against a plain scan over pseudo-random masks: returns the
// number of disagreements
extern "C" int select_bit_mismatches(int trials) {
  unsigned long long x = 88172645463325252ull;
  int bad = 0;
  for (int t = 0; t < trials; t++) {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    const unsigned long long m = (t % 3 == 0) ? (x & (x >> 11)) : x;  // sparser masks too
    int r = 0;
    for (int b = 0; b < 64; b++)
      if ((m >> b) & 1ull) {
        if (spk::dc_select_bit(m, r) != b) bad++;
        r++;
      }
  }
  return bad;
}

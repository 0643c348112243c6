// Drives the host-side table builders (chns_tables.cpp) and the host-only
// ABI entry points on exactly-sized heap buffers, for every model, a range
// of sizes and the optional-physics switches.  Built by
// tests/test_host_sanitizers.py with g++ -fsanitize=address,undefined: an
// out-of-bounds write, a use of uninitialised memory in the builders' index
// arithmetic or undefined behaviour aborts the run.  Exit status 0 and the
// line "ok <cases>" mean every case ran and every table entry is finite.
#include <cmath>
#include <cstdio>
#include <vector>

#include "chns.h"
#include "chns_host.hpp"

namespace {

int bad = 0;

void finite(const std::vector<double>& v, const char* what, int N) {
  for (double x : v)
    if (!std::isfinite(x)) {
      std::printf("non-finite %s at N=%d\n", what, N);
      ++bad;
      return;
    }
}

void run_case(int model, int N, bool vpm, bool filter) {
  chns_params p;
  chns_params_default(&p, model);
  p.Nx = p.Ny = N;
  p.lx = p.ly = 2.0 * M_PI;
  p.epsilon = 1.5 * p.lx / N;
  p.vpm = vpm ? 1 : 0;
  p.sponge = vpm ? 1 : 0;
  p.smooth_filter = filter ? 1 : 0;
  if (const char* why = chns_host::check_params(&p)) {
    std::printf("rejected N=%d: %s\n", N, why);
    ++bad;
    return;
  }
  chns_host::Derived d;
  chns_host::derive(&p, &d);
  const size_t H = (size_t)d.H, HH = H * H, NH = (size_t)N * H;
  std::vector<double> tw1(2 * HH), tc1(2 * HH), tw2(HH), tc2(HH);
  chns_host::build_coef_tables(&p, tw1.data(), tc1.data(), tw2.data(), tc2.data());
  finite(tw1, "tw1", N);
  finite(tw2, "tw2", N);
  if (model != CHNS_MODEL_NS) {
    finite(tc1, "tc1", N);
    finite(tc2, "tc2", N);
  }
  std::vector<double> tw(2 * (size_t)N);
  chns_host::build_twiddles(N, tw.data());
  finite(tw, "twiddles", N);
  std::vector<int> amax(H);
  chns_host::dealias_extent(&p, amax.data());
  std::vector<double> chi(N), spg(N);
  chns_host::build_wall_tables(&p, chi.data(), spg.data());
  finite(chi, "chi", N);
  std::vector<double> ft(HH);
  chns_host::build_filter_table(&p, ft.data());
  finite(ft, "filter", N);
  // host-only ABI entry points, in the reference layout IDXC [Nx][Ny/2+1]
  std::vector<double> eta(NH), z1w(NH), z2w(NH), beta(NH), z1c(NH), z2c(NH), ftab(NH);
  if (chns_etd_tables(&p, eta.data(), z1w.data(), z2w.data(), beta.data(), z1c.data(), z2c.data()) != CHNS_OK) {
    std::printf("chns_etd_tables failed at N=%d\n", N);
    ++bad;
  }
  if (chns_filter_table(&p, ftab.data()) != CHNS_OK || chns_dealias_extent(&p, amax.data()) != CHNS_OK) {
    std::printf("host table entry failed at N=%d\n", N);
    ++bad;
  }
  finite(eta, "eta", N);
  finite(ftab, "filter table", N);
}

}  // namespace

int main() {
  int cases = 0;
  for (int model : {CHNS_MODEL_NS, CHNS_MODEL_BINARY, CHNS_MODEL_TERNARY})
    for (int N : {16, 32, 128, 512, 2048})
      for (int sw = 0; sw < 3; ++sw) {
        run_case(model, N, sw == 1, sw == 2);
        ++cases;
      }
  // rejected parameters return an error and touch no buffer
  chns_params p;
  chns_params_default(&p, CHNS_MODEL_TERNARY);
  p.Nx = 48;
  p.Ny = 48;
  double x = 0.0;
  int a = 0;
  if (chns_etd_tables(&p, &x, &x, &x, &x, &x, &x) == CHNS_OK || chns_dealias_extent(&p, &a) == CHNS_OK) {
    std::printf("N=48 accepted\n");
    ++bad;
  }
  if (bad) return 1;
  std::printf("ok %d\n", cases);
  return 0;
}

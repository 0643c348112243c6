// Host-side internals shared by the table builder (g++) and the API (hipcc).
#pragma once
#include <cstddef>

#include "chns.h"

namespace chns_host {

struct Derived {
  int N, log2n, H;
  double dx, eps, factor_x, factor_y;
  double g1, g2, g3, del;
};

const char* check_params(const chns_params* p);
void derive(const chns_params* p, Derived* d);
// tab1: H*H entries {eta, z1w, beta, z1c}; tab2: H*H entries {z2w, z2c};
// entry (j, |i|) at j*H + |i|.
void build_coef_tables(const chns_params* p, double* tab1, double* tab2);
void build_twiddles(int N, double* tw);
void dealias_extent(const chns_params* p, int* amax);

}  // namespace chns_host

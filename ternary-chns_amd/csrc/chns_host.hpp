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
// Entry (j, |i|) at j*H + |i|: tw1 {eta, z1w} and tc1 {beta, z1c} pairs,
// tw2 z2w and tc2 z2c scalars.
void build_coef_tables(const chns_params* p, double* tw1, double* tc1, double* tw2, double* tc2);
void build_twiddles(int N, double* tw);
void dealias_extent(const chns_params* p, int* amax);
// wall profile chi[Nx] and sponge[Ny] (times 1/eta) of init.c:184-211
void build_wall_tables(const chns_params* p, double* chi, double* sponge);
// SMOOTH_FILTER factor exp(-36 pp^18) of truncation() (functions.c:54,70-73)
// at (j, |i|): ft[j*H + |i|], H = Ny/2+1
void build_filter_table(const chns_params* p, double* ft);
// 1 on the modes the cut keeps (|i| <= amax[j]), 0 elsewhere, in the layout of
// build_filter_table: the factor table of the 2/3-rule and N/3 cuts
void build_cut_table(const chns_params* p, const int* amax, double* ft);

}  // namespace chns_host

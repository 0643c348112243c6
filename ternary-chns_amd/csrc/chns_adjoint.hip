// Kernel instantiations and launch wrappers of the discrete adjoint (gfx950).
#include <hip/hip_runtime.h>

#include "chns_adjoint.hpp"
#include "chns_adjoint_launch.hpp"
#include "chns_passes.hpp"

namespace chns {

namespace {
using passes::ceil_div;
using passes::round16;

template <int L> int adj_rowD(const AdjArgs& a, hipStream_t s) {
  k_adj_rowD1<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(a);
  int e = (int)hipGetLastError();
  if (e) return e;
  k_adj_rowD2<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L> int adj_colBC(const AdjArgs& a, hipStream_t s) {
  k_adj_colBC<L><<<round16(ceil_div<L>(a.P.Jt) * 2), ColCfg<L>::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L> int adj_rowB(const AdjArgs& a, hipStream_t s) {
  k_adj_rowB<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L> int adj_colE(int stage, const AdjArgs& a, hipStream_t s) {
  const dim3 g(round16(ceil_div<L>(a.P.Jt) * 3));
  if (stage == 2) k_adj_colE<L, 2><<<g, ColCfg<L>::NT, 0, s>>>(a);
  else if (stage == 1) k_adj_colE<L, 1><<<g, ColCfg<L>::NT, 0, s>>>(a);
  else k_adj_colE<L, 0><<<g, ColCfg<L>::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L> int row_blocks_t() { return (1 << L) / RowCfg<L>::R; }
template <int L> int colBC_blocks_t(int Jt) { return round16(ceil_div<L>(Jt) * 2); }
}  // namespace

int launch_adj_rowD(int log2n, const AdjArgs& a, hipStream_t s) { CHNS_DISPATCH(log2n, (adj_rowD<L>(a, s))); }
int launch_adj_colBC(int log2n, const AdjArgs& a, hipStream_t s) { CHNS_DISPATCH(log2n, (adj_colBC<L>(a, s))); }
int launch_adj_rowB(int log2n, const AdjArgs& a, hipStream_t s) { CHNS_DISPATCH(log2n, (adj_rowB<L>(a, s))); }
int launch_adj_colE(int log2n, int stage, const AdjArgs& a, hipStream_t s) {
  CHNS_DISPATCH(log2n, (adj_colE<L>(stage, a, s)));
}
int adj_row_blocks(int log2n) { CHNS_DISPATCH(log2n, (row_blocks_t<L>())); }
int adj_colBC_blocks(int log2n, int Jt) { CHNS_DISPATCH(log2n, (colBC_blocks_t<L>(Jt))); }

int launch_adj_seed_spectrum(const cd* C1, cd* const* Ub, const double* dS, const int* amask, int N, double f,
                             double dk, int nsh, hipStream_t s) {
  const size_t n = (size_t)N * (N / 2 + 1);
  const double coef = 4.0 / ((double)N * (double)N);
  k_adj_seed_spectrum<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(C1, Ub[0], Ub[1], Ub[2], dS, amask, N, f, dk,
                                                                   coef, nsh);
  return (int)hipGetLastError();
}

int launch_fe_adjoint(double* const* f, int N, double g1, double g2, double g3, double b, double q, hipStream_t s) {
  const size_t n = (size_t)N * N;
  k_fe_adjoint<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(f[0], f[1], f[2], f[3], f[4], f[5], n, g1, g2, g3, b, q);
  return (int)hipGetLastError();
}

int launch_fe_seed(cd* const* X, cd* const* Ub, int N, double f, hipStream_t s) {
  const size_t n = (size_t)N * (N / 2 + 1);
  k_fe_seed<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(X[0], X[1], X[2], X[3], X[4], X[5], Ub[0], Ub[1], Ub[2], N, f);
  return (int)hipGetLastError();
}

int launch_adj_accum(const double* part, int nslot, size_t slot_stride, int mb, const int* nbq, double* acc,
                     hipStream_t s) {
  k_adj_accum<<<ADJ_NQ, 256, 0, s>>>(part, nslot, slot_stride, mb, nbq, acc);
  return (int)hipGetLastError();
}

}  // namespace chns

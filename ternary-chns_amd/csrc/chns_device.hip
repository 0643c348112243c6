// Kernel instantiations and launch wrappers (gfx950).
#include <hip/hip_runtime.h>

#include "chns_passes.hpp"
#include "chns_diag.hpp"
#include "chns_launch.hpp"

namespace chns {

namespace {
using passes::ceil_div;
using passes::round16;

template <int L>
int row_r2c(const double* a, const double* b, cd* oa, cd* ob, const cd* tw, int ldF, hipStream_t s) {
  k_row_r2c<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(a, b, oa, ob, tw, ldF);
  return (int)hipGetLastError();
}
template <int L>
int col_fwd(bool trunc, const cd* in, cd* out, const cd* tw, const int* amask, const double* ft, int ldF,
            hipStream_t s) {
  const dim3 g(ceil_div<L, 16>((1 << L) / 2 + 1));
  if (trunc) k_col_fwd<L, true><<<g, ColCfg<L, 16>::NT, 0, s>>>(in, out, tw, amask, ft, ldF);
  else k_col_fwd<L, false><<<g, ColCfg<L, 16>::NT, 0, s>>>(in, out, tw, amask, ft, ldF);
  return (int)hipGetLastError();
}
}  // namespace

int launch_colA(int log2n, const Bufs& B, const KParams& P, int nph, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colA<L>(Inst{B, P}, P.Jt, nph, 1, s)));
}
int launch_rowB(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowB<L>(model, Inst{B, P}, 1, s)));
}
int launch_colC(int log2n, int model, const Bufs& B, const KParams& P, bool cons, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colC<L>(model, Inst{B, P}, P.Jt, 1, s, cons)));
}
int launch_rowD(int log2n, int model, const Bufs& B, const KParams& P, bool cons, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowD<L>(model, Inst{B, P}, 1, s, cons)));
}
int launch_colE(int log2n, int model, int stage, const Bufs& B, const KParams& P, int nf, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colE<L>(model, stage, Inst{B, P}, nf, 1, s)));
}
int launch_row_r2c(int log2n, const double* a, const double* b, cd* oa, cd* ob, const cd* tw, int ldF,
                   hipStream_t s) {
  CHNS_DISPATCH(log2n, (row_r2c<L>(a, b, oa, ob, tw, ldF, s)));
}
int launch_col_fwd(int log2n, bool trunc, const cd* in, cd* out, const cd* tw, const int* amask, const double* ft,
                   int ldF, hipStream_t s) {
  CHNS_DISPATCH(log2n, (col_fwd<L>(trunc, in, out, tw, amask, ft, ldF, s)));
}
int launch_truncate(cd* S, const int* amask, const double* ft, int N, hipStream_t s) {
  const size_t n = (size_t)N * (N / 2 + 1);
  k_truncate<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(S, amask, ft, N, N / 2 + 1);
  return (int)hipGetLastError();
}
int launch_gravity(const cd* C1, const cd* C2, cd* GR, double gr1, double gr2, double f, int N, hipStream_t s) {
  const size_t n = (size_t)N * (N / 2 + 1);
  k_gravity<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(C1, C2, GR, gr1, gr2, f, N);
  return (int)hipGetLastError();
}
int launch_transpose(const cd* in, cd* out, int rows, int cols, hipStream_t s) {
  const dim3 g((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  k_transpose<<<g, 256, 0, s>>>(in, out, rows, cols);
  return (int)hipGetLastError();
}
int col_tile(int log2n) { CHNS_DISPATCH(log2n, (ColCfg<L>::C)); }

namespace {
template <int L, int OP, bool TR>
int col_inv_full_t(const cd* S, cd* out, const cd* tw, const int* amask, const double* ft, int ldF, double f,
                   hipStream_t s) {
  const dim3 g(round16(ceil_div<L>((1 << L) / 2 + 1)));
  k_col_inv_full<L, OP, TR><<<g, ColCfg<L>::NT, 0, s>>>(S, out, tw, amask, ft, ldF, f);
  return (int)hipGetLastError();
}
template <int L>
int col_inv_full(int op, bool tr, const cd* S, cd* out, const cd* tw, const int* amask, const double* ft, int ldF,
                 double f, hipStream_t s) {
#define CHNS_OPC(O)                                                                \
  if (op == O) return tr ? col_inv_full_t<L, O, true>(S, out, tw, amask, ft, ldF, f, s) \
                         : col_inv_full_t<L, O, false>(S, out, tw, amask, ft, ldF, f, s);
  CHNS_OPC(COP_ID) CHNS_OPC(COP_IKX) CHNS_OPC(COP_IKY) CHNS_OPC(COP_U) CHNS_OPC(COP_V)
#undef CHNS_OPC
  return (int)hipErrorInvalidValue;
}
template <int L>
int row_c2r_full(const cd* X, const cd* Y, double* oa, double* ob, const cd* tw, int ldF, double scale,
                 hipStream_t s) {
  k_row_c2r_full<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(X, Y, oa, ob, tw, ldF, scale);
  return (int)hipGetLastError();
}
}  // namespace

int launch_col_inv_full(int log2n, int op, bool trunc, const cd* S, cd* out, const cd* tw, const int* amask,
                        const double* ft, int ldF, double f, hipStream_t s) {
  CHNS_DISPATCH(log2n, (col_inv_full<L>(op, trunc, S, out, tw, amask, ft, ldF, f, s)));
}
int launch_row_c2r_full(int log2n, const cd* X, const cd* Y, double* oa, double* ob, const cd* tw, int ldF,
                        double scale, hipStream_t s) {
  CHNS_DISPATCH(log2n, (row_c2r_full<L>(X, Y, oa, ob, tw, ldF, scale, s)));
}
int launch_reduce(int kind, const RedArgs& A, double* partial, double* out, hipStream_t s) {
  const size_t np = (size_t)A.N * A.N;
  const int nb = (int)((np + RB_THREADS * RB_PTS - 1) / (RB_THREADS * RB_PTS));
  switch (kind) {
    case RED_FE_TER: k_reduce<RED_FE_TER><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    case RED_FE_BIN: k_reduce<RED_FE_BIN><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    case RED_KE_TER: k_reduce<RED_KE_TER><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    case RED_KE_BIN: k_reduce<RED_KE_BIN><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    case RED_KE_NS: k_reduce<RED_KE_NS><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    case RED_PERIM: k_reduce<RED_PERIM><<<nb, RB_THREADS, 0, s>>>(A, partial); break;
    default: return (int)hipErrorInvalidValue;
  }
  int e = (int)hipGetLastError();
  if (e) return e;
  k_reduce_final<<<1, RB_THREADS, 0, s>>>(partial, nb, out);
  return (int)hipGetLastError();
}
int reduce_blocks(int N) { return (int)(((size_t)N * N + RB_THREADS * RB_PTS - 1) / (RB_THREADS * RB_PTS)); }
int launch_spectrum(const SpecArgs& A, double* out, double* outy, double* outx, hipStream_t s) {
  k_spec_rows<<<(A.N + 63) / 64, 64, 0, s>>>(A);
  int e = (int)hipGetLastError();
  if (e) return e;
  k_spec_sum<<<A.nsh + A.nxy, 256, 0, s>>>(A, out, outy, outx);
  return (int)hipGetLastError();
}

}  // namespace chns

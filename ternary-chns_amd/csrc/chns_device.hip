// Kernel instantiations and launch wrappers (gfx950).
#include <hip/hip_runtime.h>

#include "chns_kernels.hpp"
#include "chns_diag.hpp"
#include "chns_launch.hpp"

namespace chns {

namespace {
template <int L, int PP = 8> constexpr int ceil_div(int a) { return (a + ColCfg<L, PP>::C - 1) / ColCfg<L, PP>::C; }

int round16(int n) { return (n + 15) / 16 * 16; }
int groupsC(int model) { return model == MODEL_TERNARY ? 5 : (model == MODEL_BINARY ? 4 : 2); }

template <int L> int colA(const Bufs& B, const KParams& P, int nph, hipStream_t s) {
  k_colA<L><<<dim3(ceil_div<L, 16>(P.Jt), nph), ColCfg<L, 16>::NT, 0, s>>>(B, P);
  return (int)hipGetLastError();
}
template <int L> int rowB(int model, const Bufs& B, const KParams& P, hipStream_t s) {
  const dim3 g(P.N / RowCfg<L>::R);
  if (model == MODEL_TERNARY) k_rowB<L, MODEL_TERNARY><<<g, RowCfg<L>::NT, 0, s>>>(B, P);
  else k_rowB<L, MODEL_BINARY><<<g, RowCfg<L>::NT, 0, s>>>(B, P);
  return (int)hipGetLastError();
}
template <int L> int colC(int model, const Bufs& B, const KParams& P, hipStream_t s) {
  const dim3 g(round16(ceil_div<L>(P.Jt) * groupsC(model)));
  if (model == MODEL_TERNARY) k_colC<L, MODEL_TERNARY><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  else if (model == MODEL_BINARY) k_colC<L, MODEL_BINARY><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  else k_colC<L, MODEL_NS><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  return (int)hipGetLastError();
}
template <int L> int rowD(int model, const Bufs& B, const KParams& P, hipStream_t s) {
  const dim3 g(P.N / RowCfg<L>::R);
  if (model == MODEL_TERNARY) k_rowD<L, MODEL_TERNARY><<<g, RowCfg<L>::NT, 0, s>>>(B, P);
  else if (model == MODEL_BINARY) k_rowD<L, MODEL_BINARY><<<g, RowCfg<L>::NT, 0, s>>>(B, P);
  else k_rowD<L, MODEL_NS><<<g, RowCfg<L>::NT, 0, s>>>(B, P);
  return (int)hipGetLastError();
}
template <int L, int ST> int colE_st(int model, const Bufs& B, const KParams& P, int nf, hipStream_t s) {
  const dim3 g(round16(ceil_div<L>(P.N / 2 + 1) * nf));
  if (model == MODEL_TERNARY) k_colE<L, MODEL_TERNARY, ST><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  else if (model == MODEL_BINARY) k_colE<L, MODEL_BINARY, ST><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  else k_colE<L, MODEL_NS, ST><<<g, ColCfg<L>::NT, 0, s>>>(B, P);
  return (int)hipGetLastError();
}
template <int L> int colE(int model, int stage, const Bufs& B, const KParams& P, int nf, hipStream_t s) {
  return stage == 1 ? colE_st<L, 1>(model, B, P, nf, s) : colE_st<L, 2>(model, B, P, nf, s);
}
template <int L>
int row_r2c(const double* a, const double* b, cd* oa, cd* ob, const cd* tw, int ldF, hipStream_t s) {
  k_row_r2c<L><<<(1 << L) / RowCfg<L>::R, RowCfg<L>::NT, 0, s>>>(a, b, oa, ob, tw, ldF);
  return (int)hipGetLastError();
}
template <int L>
int col_fwd(bool trunc, const cd* in, cd* out, const cd* tw, const int* amask, int ldF, hipStream_t s) {
  const dim3 g(ceil_div<L, 16>((1 << L) / 2 + 1));
  if (trunc) k_col_fwd<L, true><<<g, ColCfg<L, 16>::NT, 0, s>>>(in, out, tw, amask, ldF);
  else k_col_fwd<L, false><<<g, ColCfg<L, 16>::NT, 0, s>>>(in, out, tw, amask, ldF);
  return (int)hipGetLastError();
}
}  // namespace

#define CHNS_DISPATCH(log2n, CALL)          \
  switch (log2n) {                          \
    case 4: { constexpr int L = 4; return CALL; }   \
    case 5: { constexpr int L = 5; return CALL; }   \
    case 6: { constexpr int L = 6; return CALL; }   \
    case 7: { constexpr int L = 7; return CALL; }   \
    case 8: { constexpr int L = 8; return CALL; }   \
    case 9: { constexpr int L = 9; return CALL; }   \
    case 10: { constexpr int L = 10; return CALL; } \
    case 11: { constexpr int L = 11; return CALL; } \
    default: return (int)hipErrorInvalidValue;      \
  }

int launch_colA(int log2n, const Bufs& B, const KParams& P, int nph, hipStream_t s) {
  CHNS_DISPATCH(log2n, (colA<L>(B, P, nph, s)));
}
int launch_rowB(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (rowB<L>(model, B, P, s)));
}
int launch_colC(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (colC<L>(model, B, P, s)));
}
int launch_rowD(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (rowD<L>(model, B, P, s)));
}
int launch_colE(int log2n, int model, int stage, const Bufs& B, const KParams& P, int nf, hipStream_t s) {
  CHNS_DISPATCH(log2n, (colE<L>(model, stage, B, P, nf, s)));
}
int launch_row_r2c(int log2n, const double* a, const double* b, cd* oa, cd* ob, const cd* tw, int ldF,
                   hipStream_t s) {
  CHNS_DISPATCH(log2n, (row_r2c<L>(a, b, oa, ob, tw, ldF, s)));
}
int launch_col_fwd(int log2n, bool trunc, const cd* in, cd* out, const cd* tw, const int* amask, int ldF,
                   hipStream_t s) {
  CHNS_DISPATCH(log2n, (col_fwd<L>(trunc, in, out, tw, amask, ldF, s)));
}
int launch_truncate(cd* S, const int* amask, int N, hipStream_t s) {
  const size_t n = (size_t)N * (N / 2 + 1);
  k_truncate<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(S, amask, N, N / 2 + 1);
  return (int)hipGetLastError();
}
int col_tile(int log2n) { CHNS_DISPATCH(log2n, (ColCfg<L>::C)); }

namespace {
template <int L, int OP, bool TR>
int col_inv_full_t(const cd* S, cd* out, const cd* tw, const int* amask, int ldF, double f, hipStream_t s) {
  const dim3 g(round16(ceil_div<L>((1 << L) / 2 + 1)));
  k_col_inv_full<L, OP, TR><<<g, ColCfg<L>::NT, 0, s>>>(S, out, tw, amask, ldF, f);
  return (int)hipGetLastError();
}
template <int L>
int col_inv_full(int op, bool tr, const cd* S, cd* out, const cd* tw, const int* amask, int ldF, double f,
                 hipStream_t s) {
#define CHNS_OPC(O)                                                          \
  if (op == O) return tr ? col_inv_full_t<L, O, true>(S, out, tw, amask, ldF, f, s) \
                         : col_inv_full_t<L, O, false>(S, out, tw, amask, ldF, f, s);
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
                        int ldF, double f, hipStream_t s) {
  CHNS_DISPATCH(log2n, (col_inv_full<L>(op, trunc, S, out, tw, amask, ldF, f, s)));
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
  const int n = A.nsh > A.nxy ? A.nsh : A.nxy;
  k_spec_sum<<<(n + 63) / 64, 64, 0, s>>>(A, out, outy, outx);
  return (int)hipGetLastError();
}

}  // namespace chns

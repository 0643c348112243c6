// Kernel instantiations and launch wrappers (gfx950).
#include <hip/hip_runtime.h>

#include "chns_kernels.hpp"
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

}  // namespace chns

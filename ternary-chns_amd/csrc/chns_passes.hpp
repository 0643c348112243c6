// Launch shapes of the five pass kernels, shared by the single-instance
// launchers (chns_device.hip) and the batched ensemble launchers
// (chns_batch.hip).  `a` is an Inst (nb = 1) or an InstBatch of nb instances
// of one grid size and model; grid.z selects the instance.
#pragma once
#include <hip/hip_runtime.h>

#include "chns_kernels.hpp"

namespace chns {
namespace passes {

template <int L, int PP = 8> constexpr int ceil_div(int a) { return (a + ColCfg<L, PP>::C - 1) / ColCfg<L, PP>::C; }
inline int round16(int n) { return (n + 15) / 16 * 16; }
inline int groupsC(int model) { return model == MODEL_TERNARY ? 6 : (model == MODEL_BINARY ? 4 : 2); }

template <int L, bool FILT = false, class A> int colA(const A& a, int Jt, int nph, int nb, hipStream_t s) {
  k_colA<L, A, FILT><<<dim3(ceil_div<L, 16>(Jt), nph, nb), ColCfg<L, 16>::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L, bool FILT = false, int RV = 0, class A> int rowB(int model, const A& a, int nb, hipStream_t s) {
  using RC = RowCfg<L, RV>;
  if constexpr (L == 11 && RV == 0 && !FILT) {
    // 768 workgroups (3 per CU: 140 VGPRs) taking 2 or 3 rows in turn, each
    // row's loads under the previous row's transforms (k_rowB_pipe)
    const dim3 g((1 << L) / 8 * 3, 1, nb);
    if (model == MODEL_TERNARY) k_rowB_pipe<L, MODEL_TERNARY, A><<<g, RC::NT, 0, s>>>(a);
    else k_rowB_pipe<L, MODEL_BINARY, A><<<g, RC::NT, 0, s>>>(a);
    return (int)hipGetLastError();
  }
  const dim3 g((1 << L) / RC::R, 1, nb);
  if (model == MODEL_TERNARY) k_rowB<L, MODEL_TERNARY, A, FILT, RV><<<g, RC::NT, 0, s>>>(a);
  else k_rowB<L, MODEL_BINARY, A, FILT, RV><<<g, RC::NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
// cons: the conservative vorticity advection of passes C and D (binary and
// ternary with the pruned cut; NS and the unpruned passes ignore it).  Passes
// C and D of one stage must agree on it.
template <int L, bool FILT = false, int PP = 8, class A>
int colC(int model, const A& a, int Jt, int nb, hipStream_t s, bool cons = false) {
  const dim3 g(round16(ceil_div<L, PP>(Jt) * groupsC(model)), 1, nb);
  constexpr int NT = ColCfg<L, PP>::NT;
  constexpr bool CF = !FILT;   // the conservative form has pruned instantiations only
  if (model == MODEL_TERNARY) {
    if (CF && cons) k_colC<L, MODEL_TERNARY, A, FILT, PP, CF><<<g, NT, 0, s>>>(a);
    else k_colC<L, MODEL_TERNARY, A, FILT, PP><<<g, NT, 0, s>>>(a);
  } else if (model == MODEL_BINARY) {
    if (CF && cons) k_colC<L, MODEL_BINARY, A, FILT, PP, CF><<<g, NT, 0, s>>>(a);
    else k_colC<L, MODEL_BINARY, A, FILT, PP><<<g, NT, 0, s>>>(a);
  } else {
    k_colC<L, MODEL_NS, A, FILT, PP><<<g, NT, 0, s>>>(a);
  }
  return (int)hipGetLastError();
}
template <int L, bool FILT = false, int RV = 0, class A>
int rowD(int model, const A& a, int nb, hipStream_t s, bool cons = false) {
  using RC = RowCfg<L, RV>;
  const dim3 g((1 << L) / RC::R, 1, nb);
  constexpr bool CF = !FILT;
  if (model == MODEL_TERNARY) {
    if (CF && cons) k_rowD<L, MODEL_TERNARY, A, FILT, RV, CF><<<g, RC::NT, 0, s>>>(a);
    else k_rowD<L, MODEL_TERNARY, A, FILT, RV><<<g, RC::NT, 0, s>>>(a);
  } else if (model == MODEL_BINARY) {
    if (CF && cons) k_rowD<L, MODEL_BINARY, A, FILT, RV, CF><<<g, RC::NT, 0, s>>>(a);
    else k_rowD<L, MODEL_BINARY, A, FILT, RV><<<g, RC::NT, 0, s>>>(a);
  } else {
    k_rowD<L, MODEL_NS, A, FILT, RV><<<g, RC::NT, 0, s>>>(a);
  }
  return (int)hipGetLastError();
}
template <int L, int ST, bool FILT = false, int PP = 8, class A>
int colE_st(int model, const A& a, int nf, int nb, hipStream_t s) {
  // omega tiles (rounded up to an even count), then nf - 1 phase items per tile (k_colE)
  const int ntl = ceil_div<L, PP>((1 << L) / 2 + 1);
  const dim3 g(round16(((ntl + 1) & ~1) + (nf - 1) * ntl), 1, nb);
  constexpr int NT = ColCfg<L, PP>::NT;
  if (model == MODEL_TERNARY) k_colE<L, MODEL_TERNARY, ST, A, FILT, PP><<<g, NT, 0, s>>>(a);
  else if (model == MODEL_BINARY) k_colE<L, MODEL_BINARY, ST, A, FILT, PP><<<g, NT, 0, s>>>(a);
  else k_colE<L, MODEL_NS, ST, A, FILT, PP><<<g, NT, 0, s>>>(a);
  return (int)hipGetLastError();
}
template <int L, bool FILT = false, int PP = 8, class A>
int colE(int model, int stage, const A& a, int nf, int nb, hipStream_t s) {
  return stage == 1 ? colE_st<L, 1, FILT, PP>(model, a, nf, nb, s) : colE_st<L, 2, FILT, PP>(model, a, nf, nb, s);
}

}  // namespace passes

#define CHNS_DISPATCH(log2n, CALL)                  \
  switch (log2n) {                                  \
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

}  // namespace chns

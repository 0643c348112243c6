// The FFT-schedule variants of the batched passes (chns_batch.hip): column
// passes C and E with 16 points per thread (radix-16 stages: one exchange
// through LDS fewer per 2048-point transform, twice the registers per
// thread), row passes B and D with half the rows per workgroup.  At N < 256
// the column variant equals the default (8 points per thread either way).
#include <hip/hip_runtime.h>

#include "chns_launch.hpp"
#include "chns_passes.hpp"

namespace chns {

int launch_batch_rowB_half(int log2n, int model, const Bufs* B, const KParams* P, int nb, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowB<L, false, 1>(model, InstBatch{B, P}, nb, s)));
}
int launch_batch_rowD_half(int log2n, int model, const Bufs* B, const KParams* P, int nb, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowD<L, false, 1>(model, InstBatch{B, P}, nb, s, true)));
}
int launch_batch_colC_r16(int log2n, int model, const Bufs* B, const KParams* P, int Jt, int nb, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colC<L, false, 16>(model, InstBatch{B, P}, Jt, nb, s, true)));
}
int launch_batch_colE_r16(int log2n, int model, int stage, const Bufs* B, const KParams* P, int nf, int nb,
                          hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colE<L, false, 16>(model, stage, InstBatch{B, P}, nf, nb, s)));
}

}  // namespace chns

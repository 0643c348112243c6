// Batched pass launchers for ensembles (gfx950): nb instances of one grid size
// and model per launch, instance b's buffers and parameters at B[b], P[b]
// (device arrays), grid.z = b.  `sched` selects the FFT schedule of the
// group (chosen by chns_ensemble_tune): bit 0 runs the column passes C and E
// with 16 points per thread (radix-16 stages) instead of 8, bit 1 gives the
// row passes B and D half the rows per workgroup; those variants are
// instantiated in chns_batch_sched.hip.  Passes C and D take the
// conservative vorticity advection (ensembles have no sponge).
#include <hip/hip_runtime.h>

#include "chns_launch.hpp"
#include "chns_passes.hpp"

namespace chns {

int launch_batch_colA(int log2n, const Bufs* B, const KParams* P, int Jt, int nph, int nb, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colA<L>(InstBatch{B, P}, Jt, nph, nb, s)));
}
int launch_batch_rowB(int log2n, int model, const Bufs* B, const KParams* P, int nb, int sched, hipStream_t s) {
  if (sched & 2) return launch_batch_rowB_half(log2n, model, B, P, nb, s);
  CHNS_DISPATCH(log2n, (passes::rowB<L>(model, InstBatch{B, P}, nb, s)));
}
int launch_batch_colC(int log2n, int model, const Bufs* B, const KParams* P, int Jt, int nb, int sched,
                      hipStream_t s) {
  if (sched & 1) return launch_batch_colC_r16(log2n, model, B, P, Jt, nb, s);
  CHNS_DISPATCH(log2n, (passes::colC<L>(model, InstBatch{B, P}, Jt, nb, s, true)));
}
int launch_batch_rowD(int log2n, int model, const Bufs* B, const KParams* P, int nb, int sched, hipStream_t s) {
  if (sched & 2) return launch_batch_rowD_half(log2n, model, B, P, nb, s);
  CHNS_DISPATCH(log2n, (passes::rowD<L>(model, InstBatch{B, P}, nb, s, true)));
}
int launch_batch_colE(int log2n, int model, int stage, const Bufs* B, const KParams* P, int nf, int nb, int sched,
                      hipStream_t s) {
  if (sched & 1) return launch_batch_colE_r16(log2n, model, stage, B, P, nf, nb, s);
  CHNS_DISPATCH(log2n, (passes::colE<L>(model, stage, InstBatch{B, P}, nf, nb, s)));
}

}  // namespace chns

// Instantiations of the five pass kernels with the SMOOTH_FILTER form of
// truncation() (macro.h:15, functions.c:70-73): every mode is kept and
// multiplied by exp(-36 pp^18) twice wherever the reference truncates, so the
// transforms are not pruned and pass D loads its rows without the stager.
// A translation unit of its own keeps the build of chns_device.hip short.
#include <hip/hip_runtime.h>

#include "chns_passes.hpp"
#include "chns_launch.hpp"

namespace chns {

int launch_filt_colA(int log2n, const Bufs& B, const KParams& P, int nph, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colA<L, true>(Inst{B, P}, P.Jt, nph, 1, s)));
}
int launch_filt_rowB(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowB<L, true>(model, Inst{B, P}, 1, s)));
}
int launch_filt_colC(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colC<L, true>(model, Inst{B, P}, P.Jt, 1, s)));
}
int launch_filt_rowD(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::rowD<L, true>(model, Inst{B, P}, 1, s)));
}
int launch_filt_colE(int log2n, int model, int stage, const Bufs& B, const KParams& P, int nf, hipStream_t s) {
  CHNS_DISPATCH(log2n, (passes::colE<L, true>(model, stage, Inst{B, P}, nf, 1, s)));
}

}  // namespace chns

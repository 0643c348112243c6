// Launch wrappers of the discrete-adjoint kernels (chns_adjoint.hip), called
// by the host API (chns_adjoint.cpp).  Each returns the hipError_t of the launch as int.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace chns {
struct AdjArgs;
struct cd;

int launch_adj_rowD(int log2n, const AdjArgs& a, hipStream_t s);
int launch_adj_colBC(int log2n, const AdjArgs& a, hipStream_t s);
int launch_adj_rowB(int log2n, const AdjArgs& a, hipStream_t s);
// stage 2: product at a; 1: product at u; 0: seed only
int launch_adj_colE(int log2n, int stage, const AdjArgs& a, hipStream_t s);
int launch_adj_seed_spectrum(const cd* C1, cd* const* Ub, const double* dS, const int* amask, int N, double f,
                             double dk, int nsh, hipStream_t s);
// free-energy loss: real-space adjoints in place over f[0..5] (c1, c2, grad c1,
// grad c2), then the spectral seed from their r2c X[0..5]
int launch_fe_adjoint(double* const* f, int N, double g1, double g2, double g3, double b, double q, hipStream_t s);
int launch_fe_seed(cd* const* X, cd* const* Ub, int N, double f, hipStream_t s);
int launch_adj_accum(const double* part, int nslot, size_t slot_stride, int mb, const int* nbq, double* acc,
                     hipStream_t s);
// workgroup counts of the partial-writing kernels
int adj_row_blocks(int log2n);
int adj_colBC_blocks(int log2n, int Jt);
}  // namespace chns

// Launch wrappers between the host API (chns_api.cpp, chns_ensemble.cpp) and
// the kernels (chns_device.hip, chns_batch.hip).  Each returns the hipError_t of the launch as int.
#pragma once
#include <hip/hip_runtime_api.h>

namespace chns {
struct Bufs;
struct KParams;
struct cd;

int launch_colA(int log2n, const Bufs& B, const KParams& P, int nphase, hipStream_t s);
int launch_rowB(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s);
// cons: the conservative vorticity advection (pass C writes Hw for pass D
// instead of Hwx, Hwy); the same value for passes C and D of a stage
int launch_colC(int log2n, int model, const Bufs& B, const KParams& P, bool cons, hipStream_t s);
int launch_rowD(int log2n, int model, const Bufs& B, const KParams& P, bool cons, hipStream_t s);
int launch_colE(int log2n, int model, int stage, const Bufs& B, const KParams& P, int nfields, hipStream_t s);
int launch_row_r2c(int log2n, const double* a, const double* b, cd* oa, cd* ob, const cd* tw, int ldF,
                   hipStream_t s);
// trunc: truncation() on the way, the sharp cut (amask) or, with ft, the smooth filter
int launch_col_fwd(int log2n, bool trunc, const cd* in, cd* out, const cd* tw, const int* amask, const double* ft,
                   int ldF, hipStream_t s);
int launch_truncate(cd* S, const int* amask, const double* ft, int N, hipStream_t s);
int launch_gravity(const cd* C1, const cd* C2, cd* GR, double gr1, double gr2, double f, int N, hipStream_t s);
// out[c][r] = in[r][c], row-major [rows][cols] complex in, [cols][rows] out
int launch_transpose(const cd* in, cd* out, int rows, int cols, hipStream_t s);
int col_tile(int log2n);   // columns per column-kernel workgroup

// the passes with the SMOOTH_FILTER form of truncation() (chns_filter.hip;
// single instances, every mode kept, Bufs::filt set)
int launch_filt_colA(int log2n, const Bufs& B, const KParams& P, int nphase, hipStream_t s);
int launch_filt_rowB(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s);
int launch_filt_colC(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s);
int launch_filt_rowD(int log2n, int model, const Bufs& B, const KParams& P, hipStream_t s);
int launch_filt_colE(int log2n, int model, int stage, const Bufs& B, const KParams& P, int nfields, hipStream_t s);

// batched passes (chns_batch.hip): nb instances of one size and model, device
// arrays B[nb], P[nb]; sched: the group's FFT schedule (bit 0: radix-16
// column passes, bit 1: half the rows per row workgroup).  Passes C and D
// take the conservative vorticity advection (ensembles have no sponge).
int launch_batch_colA(int log2n, const Bufs* B, const KParams* P, int Jt, int nph, int nb, hipStream_t s);
int launch_batch_rowB(int log2n, int model, const Bufs* B, const KParams* P, int nb, int sched, hipStream_t s);
int launch_batch_colC(int log2n, int model, const Bufs* B, const KParams* P, int Jt, int nb, int sched,
                      hipStream_t s);
int launch_batch_rowD(int log2n, int model, const Bufs* B, const KParams* P, int nb, int sched, hipStream_t s);
int launch_batch_colE(int log2n, int model, int stage, const Bufs* B, const KParams* P, int nf, int nb, int sched,
                      hipStream_t s);
// the schedule variants (chns_batch_sched.hip)
int launch_batch_rowB_half(int log2n, int model, const Bufs* B, const KParams* P, int nb, hipStream_t s);
int launch_batch_rowD_half(int log2n, int model, const Bufs* B, const KParams* P, int nb, hipStream_t s);
int launch_batch_colC_r16(int log2n, int model, const Bufs* B, const KParams* P, int Jt, int nb, hipStream_t s);
int launch_batch_colE_r16(int log2n, int model, int stage, const Bufs* B, const KParams* P, int nf, int nb,
                          hipStream_t s);

// diagnostics (chns_diag.hpp)
struct RedArgs;
struct SpecArgs;
int launch_col_inv_full(int log2n, int op, bool trunc, const cd* S, cd* out, const cd* tw, const int* amask,
                        const double* ft, int ldF, double f, hipStream_t s);
int launch_row_c2r_full(int log2n, const cd* X, const cd* Y, double* oa, double* ob, const cd* tw, int ldF,
                        double scale, hipStream_t s);
int launch_reduce(int kind, const RedArgs& A, double* partial, double* out, hipStream_t s);
int reduce_blocks(int N);
int launch_spectrum(const SpecArgs& A, double* out, double* outy, double* outx, hipStream_t s);
}  // namespace chns

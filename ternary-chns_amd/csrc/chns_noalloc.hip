// chns_free_energy_noalloc: the free energy of chns_free_energy
// (free_energy.c:13-157) for callers that cannot allocate device memory.
//
// chns_free_energy allocates, on its first call, two [N][N/2+1] column-iFFT
// intermediates and its reduction partials.  This variant allocates nothing:
//  * the reduction partials live in `scratch`, a static device array of
//    kScratchDoubles doubles (41 KB, enough for N <= 2048), one copy per GPU
//    for as long as the process runs; larger N is rejected;
//  * the intermediates reuse the context's stage-1 buffers NW and NC1 (the
//    saved a - z2 N(u) of a step), which hold nothing between steps, with a
//    leading dimension of N/2+1 instead of the padded ldF (same arithmetic, so
//    the results equal chns_free_energy's bit for bit).
// Nothing in the array needs clearing: every reduction block overwrites its
// own partial and the final sum overwrites the result slots.  All calls on
// one GPU share its array (the last call's partials stay in it), so calls that
// overlap on one GPU, from two threads or two streams, race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "chns.h"
#include "chns_ctx.hpp"
#include "chns_launch.hpp"

namespace {

// (reduction blocks of N = 2048: 2048^2 / 4096) partials of RED_Q sums each,
// then RED_Q result slots
constexpr int kMaxN = 2048;
constexpr size_t kScratchDoubles = ((size_t)kMaxN * kMaxN / 4096 + 1) * chns::RED_Q;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};   // device address of `scratch` per GPU

}  // namespace

// The device address of `scratch` on `dev`, which must be current.  Called by
// chns_create for its device (and again, as a no-op, by every call below).
int chns_noalloc_setup(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_ptr[dev] = (double*)p;
  }
  if (out) *out = g_ptr[dev];
  return 0;
}

extern "C" int chns_free_energy_noalloc(chns_ctx* c, double* ephi, double* egphi) {
  if (!c) return CHNS_E_ARG;
  if (!c->have_state) return chns_fail(c, CHNS_E_STATE, "no state set");
  if (c->nph == 0) return chns_fail(c, CHNS_E_ARG, "free_energy needs a phase field");
  const int N = c->d.N, H = c->d.H, L = c->d.log2n;
  const size_t need = ((size_t)chns::reduce_blocks(N) + 1) * chns::RED_Q;
  if (N > kMaxN || need > kScratchDoubles)
    return chns_fail(c, CHNS_E_ARG, "chns_free_energy_noalloc: N above the scratch maximum of 2048");
  hipError_t e = hipSetDevice(c->dev);
  if (e != hipSuccess) return chns_hip_fail(c, e, "hipSetDevice");
  double* scr = nullptr;
  int r = chns_noalloc_setup(c->dev, &scr);
  if (r) return chns_hip_fail(c, (hipError_t)r, "hipGetSymbolAddress(scratch)");
  const double inv = 1.0 / ((double)c->p.Nx * c->p.Ny);
  chns::cd* t1 = c->b.NW;     // [N][H] TJ-tiled intermediates (N*H entries each)
  chns::cd* t2 = c->b.NC1;
  double* rb[6] = {(double*)c->b.Mh1, (double*)c->b.Mh2, (double*)c->b.G1,
                   (double*)c->b.G2, (double*)c->b.G3, (double*)c->b.G4};
  const chns::cd* C1 = c->b.C1;
  const chns::cd* C2 = c->b.C2;
  auto pair = [&](const chns::cd* fa, int opa, const chns::cd* fb, int opb, double* ra, double* rbb) -> int {
    int q = chns::launch_col_inv_full(L, opa, false, fa, t1, c->b.tw, c->b.amask, c->b.filt, H, c->d.factor_x,
                                      c->stream);
    if (!q && fb)
      q = chns::launch_col_inv_full(L, opb, false, fb, t2, c->b.tw, c->b.amask, c->b.filt, H, c->d.factor_x,
                                    c->stream);
    if (!q) q = chns::launch_row_c2r_full(L, t1, fb ? t2 : nullptr, ra, fb ? rbb : nullptr, c->b.tw, H, inv, c->stream);
    return q;
  };
  chns::RedArgs A{};
  int kind;
  if (c->nph == 2) {
    // untruncated c1, c2 and their gradients (free_energy.c:218-221, 301-313)
    if ((r = pair(C1, chns::COP_ID, C2, chns::COP_ID, rb[0], rb[1])) ||
        (r = pair(C1, chns::COP_IKX, C1, chns::COP_IKY, rb[2], rb[3])) ||
        (r = pair(C2, chns::COP_IKX, C2, chns::COP_IKY, rb[4], rb[5])))
      return chns_hip_fail(c, (hipError_t)r, "free energy transforms");
    for (int i = 0; i < 6; ++i) A.f[i] = rb[i];
    kind = chns::RED_FE_TER;
  } else {
    if ((r = pair(C1, chns::COP_ID, C1, chns::COP_IKX, rb[0], rb[1])) ||
        (r = pair(C1, chns::COP_IKY, nullptr, 0, rb[2], nullptr)))
      return chns_hip_fail(c, (hipError_t)r, "free energy transforms");
    for (int i = 0; i < 3; ++i) A.f[i] = rb[i];
    kind = chns::RED_FE_BIN;
  }
  A.g1 = c->d.g1;
  A.g2 = c->d.g2;
  A.g3 = c->d.g3;
  A.N = N;
  double* out = scr + (size_t)chns::reduce_blocks(N) * chns::RED_Q;
  if ((r = chns::launch_reduce(kind, A, scr, out, c->stream))) return chns_hip_fail(c, (hipError_t)r, "reduction");
  double q[chns::RED_Q];
  if ((e = hipMemcpyAsync(q, out, sizeof q, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
    return chns_hip_fail(c, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return chns_hip_fail(c, e, "hipStreamSynchronize");
  const double eps = c->d.eps;
  double e_bulk, e_grad;
  if (c->nph == 2) {
    e_bulk = q[0] * (6.0 / eps) * inv;                              // free_energy.c:246-248
    e_grad = q[1] * (0.375 * eps * inv);                            // free_energy.c:335
  } else {
    e_bulk = q[0] * (12.0 * (c->p.sigma / eps)) * inv;              // free_energy.c:233-236
    e_grad = q[1] * ((0.75 * c->p.sigma * eps) * inv);              // free_energy.c:290
  }
  if (ephi) *ephi = e_bulk;
  if (egphi) *egphi = e_grad;
  c->energy_flag = !(e_bulk < 100.0);   // free_energy.c:151-154, reported by chns_status
  return CHNS_OK;
}

// C ABI of the device diagnostics: free_energy.c, kinetic_energy.c,
// perimeter.c, spectrum.c and the real-field dump of main.c:332-358.
// The row-pass output buffers (Mh*, G*) serve as packed real [N][N] scratch:
// they are rewritten by every step, so diagnostics between steps are free to
// use them; the dealiased row inputs (Hc*) are left untouched.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <vector>

#include "chns.h"
#include "chns_ctx.hpp"
#include "chns_launch.hpp"

namespace {

#define DHIP(ctx, call)                                                         \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) return chns_hip_fail((ctx), e_, #call);              \
  } while (0)
#define DLAUNCH(ctx, call)                                                      \
  do {                                                                          \
    int e_ = (call);                                                            \
    if (e_ != 0) return chns_hip_fail((ctx), (hipError_t)e_, #call);           \
  } while (0)

int nshell_of(int N) { return (int)std::lround(((double)N / std::sqrt(2.0)) + 1.0); }   // main.c:60

int prepare(chns_ctx* c) {
  if (!c) return CHNS_E_ARG;
  if (!c->have_state) return chns_fail(c, CHNS_E_STATE, "no state set");
  DHIP(c, hipSetDevice(c->dev));
  const size_t tiled = (size_t)c->d.N * c->ldF * sizeof(chns::cd);
  int rc;
  if (!c->dT1 && ((rc = chns_dalloc_bytes(c, (void**)&c->dT1, tiled)) ||
                  (rc = chns_dalloc_bytes(c, (void**)&c->dT2, tiled))))
    return rc;
  if (!c->dred &&
      (rc = chns_dalloc_bytes(c, (void**)&c->dred, ((size_t)chns::reduce_blocks(c->d.N) + 1) * chns::RED_Q * 8)))
    return rc;
  return 0;
}

double* real_buf(chns_ctx* c, int k) {
  chns::cd* b[6] = {c->b.Mh1, c->b.Mh2, c->b.G1, c->b.G2, c->b.G3, c->b.G4};
  return (double*)b[k];
}

const chns::cd* field(chns_ctx* c, int which) { return which == 0 ? c->b.W : (which == 1 ? c->b.C1 : c->b.C2); }

// Real-space pair (op_a on field a, op_b on field b; b < 0: none) into real buffers ra, rb.
int real_pair(chns_ctx* c, int fa, int opa, int fb, int opb, bool trunc, int ra, int rb) {
  const int L = c->d.log2n;
  DLAUNCH(c, chns::launch_col_inv_full(L, opa, trunc, field(c, fa), c->dT1, c->b.tw, c->b.amask, c->b.filt, c->ldF,
                                       c->d.factor_x, c->stream));
  if (fb >= 0)
    DLAUNCH(c, chns::launch_col_inv_full(L, opb, trunc, field(c, fb), c->dT2, c->b.tw, c->b.amask, c->b.filt, c->ldF,
                                         c->d.factor_x, c->stream));
  DLAUNCH(c, chns::launch_row_c2r_full(L, c->dT1, fb >= 0 ? c->dT2 : nullptr, real_buf(c, ra),
                                       fb >= 0 ? real_buf(c, rb) : nullptr, c->b.tw, c->ldF,
                                       1.0 / ((double)c->p.Nx * c->p.Ny), c->stream));
  return 0;
}

int reduce(chns_ctx* c, int kind, const double* const* f, double* q) {
  chns::RedArgs A{};
  for (int i = 0; i < 6; ++i) A.f[i] = f[i];
  A.g1 = c->d.g1;
  A.g2 = c->d.g2;
  A.g3 = c->d.g3;
  A.N = c->d.N;
  double* out = c->dred + (size_t)chns::reduce_blocks(c->d.N) * chns::RED_Q;
  DLAUNCH(c, chns::launch_reduce(kind, A, c->dred, out, c->stream));
  DHIP(c, hipMemcpyAsync(q, out, chns::RED_Q * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace

int chns_diag_prepare(chns_ctx* c) { return prepare(c); }

int chns_fe_fields(chns_ctx* c, bool trunc, double** f) {
  int rc;
  if (c->nph == 2) {
    // c1, c2 and their gradients (free_energy.c:218-221, 301-313)
    if ((rc = real_pair(c, 1, chns::COP_ID, 2, chns::COP_ID, trunc, 0, 1)) ||
        (rc = real_pair(c, 1, chns::COP_IKX, 1, chns::COP_IKY, trunc, 2, 3)) ||
        (rc = real_pair(c, 2, chns::COP_IKX, 2, chns::COP_IKY, trunc, 4, 5)))
      return rc;
    for (int i = 0; i < 6; ++i) f[i] = real_buf(c, i);
  } else {
    if ((rc = real_pair(c, 1, chns::COP_ID, 1, chns::COP_IKX, trunc, 0, 1)) ||
        (rc = real_pair(c, 1, chns::COP_IKY, -1, 0, trunc, 2, -1)))
      return rc;
    for (int i = 0; i < 3; ++i) f[i] = real_buf(c, i);
    for (int i = 3; i < 6; ++i) f[i] = nullptr;
  }
  return 0;
}

int chns_fe_sums(chns_ctx* c, const double* const* f, double g1, double g2, double g3, double* q) {
  chns::RedArgs A{};
  for (int i = 0; i < 6; ++i) A.f[i] = f[i];
  A.g1 = g1;
  A.g2 = g2;
  A.g3 = g3;
  A.N = c->d.N;
  double* out = c->dred + (size_t)chns::reduce_blocks(c->d.N) * chns::RED_Q;
  DLAUNCH(c, chns::launch_reduce(c->nph == 2 ? chns::RED_FE_TER : chns::RED_FE_BIN, A, c->dred, out, c->stream));
  DHIP(c, hipMemcpyAsync(q, out, chns::RED_Q * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int chns_free_energy(chns_ctx* c, double* ephi, double* egphi) {
  int rc = prepare(c);
  if (rc) return rc;
  if (c->nph == 0) return chns_fail(c, CHNS_E_ARG, "free_energy needs a phase field");
  const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);   // free_energy.c:211
  const double eps = c->d.eps;
  double q[chns::RED_Q];
  double e_bulk = 0.0, e_grad = 0.0;
  double* f[6];
  if ((rc = chns_fe_fields(c, false, f))) return rc;   // the untruncated fields
  if ((rc = chns_fe_sums(c, f, c->d.g1, c->d.g2, c->d.g3, q))) return rc;
  if (c->nph == 2) {
    e_bulk = q[0] * (6.0 / eps) * scale;                        // free_energy.c:246-248
    e_grad = q[1] * (0.375 * eps * scale);                      // free_energy.c:335
  } else {
    e_bulk = q[0] * (12.0 * (c->p.sigma / eps)) * scale;                // free_energy.c:233-236
    e_grad = q[1] * ((0.75 * c->p.sigma * eps) * scale);                // free_energy.c:290
  }
  if (ephi) *ephi = e_bulk;
  if (egphi) *egphi = e_grad;
  // the reference aborts the run here (free_energy.c:151-154); chns_status reports it
  c->energy_flag = !(e_bulk < 100.0);
  return CHNS_OK;
}

extern "C" int chns_kinetic_energy_counts(chns_ctx* c, double* eu, double* eph1, double* eph2, double* n1,
                                         double* n2) {
  int rc = prepare(c);
  if (rc) return rc;
  const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);
  // truncated u, v and phase fields (kinetic_energy.c:17-32)
  if ((rc = real_pair(c, 0, chns::COP_U, 0, chns::COP_V, true, 0, 1))) return rc;
  if (c->nph == 2 && (rc = real_pair(c, 1, chns::COP_ID, 2, chns::COP_ID, true, 2, 3))) return rc;
  if (c->nph == 1 && (rc = real_pair(c, 1, chns::COP_ID, -1, 0, true, 2, -1))) return rc;
  const double* f[6] = {real_buf(c, 0), real_buf(c, 1), c->nph >= 1 ? real_buf(c, 2) : nullptr,
                        c->nph >= 2 ? real_buf(c, 3) : nullptr, nullptr, nullptr};
  const int kind = c->nph == 2 ? chns::RED_KE_TER : (c->nph == 1 ? chns::RED_KE_BIN : chns::RED_KE_NS);
  double q[chns::RED_Q];
  if ((rc = reduce(c, kind, f, q))) return rc;
  if (eu) *eu = q[0] * scale;                                   // kinetic_energy.c:95
  if (eph1) *eph1 = q[2] != 0.0 ? q[1] / q[2] : q[1];           // kinetic_energy.c:105-107,124-127
  if (eph2) *eph2 = q[4] != 0.0 ? q[3] / q[4] : q[3];           // kinetic_energy.c:130-133
  if (n1) *n1 = c->nph >= 1 ? q[2] : 0.0;                       // count / count1
  if (n2) *n2 = c->nph == 2 ? q[4] : 0.0;                       // count2
  return CHNS_OK;
}

extern "C" int chns_kinetic_energy(chns_ctx* c, double* eu, double* eph1, double* eph2) {
  return chns_kinetic_energy_counts(c, eu, eph1, eph2, nullptr, nullptr);
}

extern "C" int chns_perimeter(chns_ctx* c, double* counts) {
  int rc = prepare(c);
  if (rc) return rc;
  if (c->nph == 0 || !counts) return chns_fail(c, CHNS_E_ARG, "perimeter needs a phase field and an output");
  // untruncated phase fields (perimeter.c:127-131)
  if (c->nph == 2 && (rc = real_pair(c, 1, chns::COP_ID, 2, chns::COP_ID, false, 0, 1))) return rc;
  if (c->nph == 1 && (rc = real_pair(c, 1, chns::COP_ID, -1, 0, false, 0, -1))) return rc;
  const double* f[6] = {real_buf(c, 0), c->nph == 2 ? real_buf(c, 1) : nullptr, nullptr, nullptr, nullptr, nullptr};
  double q[chns::RED_Q];
  if ((rc = reduce(c, chns::RED_PERIM, f, q))) return rc;
  for (int i = 0; i < 4; ++i) counts[i] = q[i];
  return CHNS_OK;
}

extern "C" int chns_spectrum(chns_ctx* c, int truncate_state, double* ek, double* ensk, double* sph1,
                             double* sph2, double* sp12, double* ekx, double* eky) {
  int rc = prepare(c);
  if (rc) return rc;
  const int N = c->d.N, nsh = nshell_of(N), nxy = N / 2 + 2;
  const size_t npart = (size_t)5 * N * nsh + (size_t)N * nxy + N;
  const size_t nout = (size_t)5 * nsh + 2 * (size_t)nxy;
  if (!c->dspec && ((rc = chns_dalloc_bytes(c, (void**)&c->dspec, npart * 8)) ||
                    (rc = chns_dalloc_bytes(c, (void**)&c->dspec_out, nout * 8))))
    return rc;
  chns::SpecArgs A{};
  A.W = c->b.W;
  A.C1 = c->b.C1;
  A.C2 = c->b.C2;
  A.amask = c->b.amask;
  A.filt = c->b.filt;
  A.f = c->d.factor_x;
  A.dk = 2.0 * M_PI / c->p.lx;                          // spectrum.c:16
  A.kymax = (c->p.Ny / 2.0) * c->d.factor_y;            // main.c:58
  const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);
  A.scale2 = scale * scale;
  A.N = N;
  A.nsh = nsh;
  A.nxy = nxy;
  A.nph = c->nph;
  A.part = c->dspec;
  A.party = c->dspec + (size_t)5 * N * nsh;   // shell-major partials, see k_spec_rows
  A.partx = A.party + (size_t)N * nxy;
  double* out = c->dspec_out;
  DLAUNCH(c, chns::launch_spectrum(A, out, out + 5 * nsh, out + 5 * nsh + nxy, c->stream));
  std::vector<double> h(nout);
  DHIP(c, hipMemcpyAsync(h.data(), out, nout * 8, hipMemcpyDeviceToHost, c->stream));
  DHIP(c, hipStreamSynchronize(c->stream));
  double* dst[5] = {ek, ensk, sph1, sph2, sp12};
  for (int q = 0; q < 5; ++q)
    if (dst[q]) std::copy(h.begin() + (size_t)q * nsh, h.begin() + (size_t)(q + 1) * nsh, dst[q]);
  if (eky) std::copy(h.begin() + 5 * nsh, h.begin() + 5 * nsh + nxy, eky);
  if (ekx) std::copy(h.begin() + 5 * nsh + nxy, h.begin() + 5 * nsh + 2 * nxy, ekx);
  if (truncate_state) return chns_truncate_state(c);     // spectrum.c:22-27
  return CHNS_OK;
}

extern "C" int chns_real_field(chns_ctx* c, int which, int truncate, double* out) {
  int rc = prepare(c);
  if (rc) return rc;
  if (!out || which < 0 || which > c->nph) return chns_fail(c, CHNS_E_ARG, "bad field index or output");
  if ((rc = real_pair(c, which, chns::COP_ID, -1, 0, truncate != 0, 0, -1))) return rc;
  const size_t nb = (size_t)c->d.N * c->d.N * sizeof(double);
  DHIP(c, hipMemcpyAsync(out, real_buf(c, 0), nb, hipMemcpyDeviceToHost, c->stream));
  DHIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

int chns_real_of(chns_ctx* c, const chns::cd* S, double* out) {
  int rc = prepare(c);
  if (rc) return rc;
  const int L = c->d.log2n;
  DLAUNCH(c, chns::launch_col_inv_full(L, chns::COP_ID, false, S, c->dT1, c->b.tw, c->b.amask, nullptr, c->ldF,
                                       c->d.factor_x, c->stream));
  DLAUNCH(c, chns::launch_row_c2r_full(L, c->dT1, nullptr, real_buf(c, 0), nullptr, c->b.tw, c->ldF,
                                       1.0 / ((double)c->p.Nx * c->p.Ny), c->stream));
  const size_t nb = (size_t)c->d.N * c->d.N * sizeof(double);
  DHIP(c, hipMemcpyAsync(out, real_buf(c, 0), nb, hipMemcpyDeviceToHost, c->stream));
  DHIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

extern "C" int chns_diag_sizes(const chns_params* p, int* nshell, int* nxy) {
  if (!p || p->Nx <= 0) return CHNS_E_ARG;
  if (nshell) *nshell = nshell_of(p->Nx);
  if (nxy) *nxy = p->Nx / 2 + 2;
  return CHNS_OK;
}

// C ABI of the discrete adjoint of the ternary ETD2RK step: forward run with
// state checkpoints, loss seeds, and the reverse sweep with per-segment
// replay (recompute) of the trajectory.  No reference counterpart: the
// reference (etd2rk.c, main.c:260-368) only integrates forward.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "chns.h"
#include "chns_adjoint_types.hpp"
#include "chns_adjoint_launch.hpp"
#include "chns_ctx.hpp"
#include "chns_launch.hpp"

struct chns_adj_state {
  int nsteps = 0, K = 0, nseg = 0;
  bool forward_done = false, seeded = false, backward_done = false;
  size_t S = 0;                  // complex entries per spectral field
  chns::cd* ckpt = nullptr;      // [nseg][3][S]: the state at steps 0, K, 2K, ...
  chns::cd* traj = nullptr;      // [2K+1][3][S]: u_0..u_K of a segment, then a_0..a_{K-1}
  int traj_K = 0, ckpt_n = 0;    // allocated capacities
  chns::cd *ub = nullptr, *ab = nullptr, *gi = nullptr;   // [3][S] each
  chns::cd* tj = nullptr;        // [6][N*ldJ]: Hxi, Hxix, Het1, Het2, Hmb1, Hmb2
  chns::cd* tf = nullptr;        // [16][N*ldJ]: O0..O13, Ocb1, Ocb2
  double* part = nullptr;        // [2][ADJ_NQ][mb]
  double* acc = nullptr;         // [ADJ_NQ]
  double* dS = nullptr;          // [nshell]
  int* nbq = nullptr;            // [ADJ_NQ]
  int mb = 0, nsh = 0;
  double dgam[3] = {0.0, 0.0, 0.0};   // direct dL/dgamma_i of the loss (the free-energy seed)
  // forward intermediates of every stage of a replayed segment (store mode):
  // slot 2k holds passes A-C of N(u_k), slot 2k+1 those of N(a_k)
  int store = 1;                 // 1: keep them when device memory allows, 0: recompute
  chns::cd* inter = nullptr;     // [2K][10 RJ + 2 RF]
  int inter_K = 0;
};

namespace {

#define AHIP(ctx, call)                                                  \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) return chns_hip_fail((ctx), e_, #call);       \
  } while (0)
#define ALAUNCH(ctx, call)                                               \
  do {                                                                   \
    int e_ = (call);                                                     \
    if (e_ != 0) return chns_hip_fail((ctx), (hipError_t)e_, #call);    \
  } while (0)

template <class T> int amalloc(chns_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) return chns_hip_fail(c, e, "hipMalloc(adjoint)");
  e = hipMemset(q, 0, count * sizeof(T));
  if (e != hipSuccess) return chns_hip_fail(c, e, "hipMemset(adjoint)");
  *p = (T*)q;
  return 0;
}
template <class T> void afree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int nshell_of(int N) { return (int)std::lround(((double)N / std::sqrt(2.0)) + 1.0); }   // main.c:60

chns::cd* field(chns::cd* base, size_t S, int k, int f) { return base + ((size_t)k * 3 + f) * S; }

// Fixed-size work buffers (allocated once per context).
int ensure_work(chns_ctx* c) {
  chns_adj_state* a = c->adj;
  if (a->ub) return 0;
  const int N = c->d.N, L = c->d.log2n;
  a->S = (size_t)N * c->d.H;
  const size_t RJ = (size_t)N * c->ldJ;
  int rc;
  if ((rc = amalloc(c, &a->ub, 3 * a->S)) || (rc = amalloc(c, &a->ab, 3 * a->S)) ||
      (rc = amalloc(c, &a->gi, 3 * a->S)) || (rc = amalloc(c, &a->tj, 6 * RJ)) ||
      (rc = amalloc(c, &a->tf, 16 * RJ)))
    return rc;
  const int nbr = chns::adj_row_blocks(L), nbc = chns::adj_colBC_blocks(L, c->Jt);
  a->mb = nbr > nbc ? nbr : nbc;
  a->nsh = nshell_of(N);
  if ((rc = amalloc(c, &a->part, 2 * (size_t)chns::ADJ_NQ * a->mb)) || (rc = amalloc(c, &a->acc, chns::ADJ_NQ)) ||
      (rc = amalloc(c, &a->dS, a->nsh)) || (rc = amalloc(c, &a->nbq, chns::ADJ_NQ)))
    return rc;
  int nb[chns::ADJ_NQ];
  for (int q = 0; q < chns::ADJ_NQ; ++q) nb[q] = nbr;
  nb[chns::AQ_MC1] = nb[chns::AQ_MC2] = nb[chns::AQ_LG1] = nb[chns::AQ_LG2] = nbc;
  AHIP(c, hipMemcpy(a->nbq, nb, sizeof(nb), hipMemcpyHostToDevice));
  return 0;
}

chns::AdjArgs adj_args(chns_ctx* c, const chns::cd* const* x, int slot) {
  chns_adj_state* a = c->adj;
  const size_t RJ = (size_t)c->d.N * c->ldJ;
  chns::AdjArgs r{};
  chns::AdjBufs& A = r.A;
  const chns::Bufs& B = c->b;
  if (x) {
    A.W = x[0];
    A.C1 = x[1];
    A.C2 = x[2];
  }
  A.Hc1 = B.Hc1; A.Hc2 = B.Hc2; A.Hu = B.Hu; A.Hv = B.Hv; A.Hwx = B.Hwx; A.Hwy = B.Hwy;
  A.Hmu1 = B.Hmu1; A.Hmu2 = B.Hmu2; A.Hdx1 = B.Hdx1; A.Hdx2 = B.Hdx2;
  A.Mh1 = B.Mh1;
  A.Mh2 = B.Mh2;
  for (int f = 0; f < 3; ++f) {
    A.Ub[f] = a->ub + f * a->S;
    A.Ab[f] = a->ab + f * a->S;
    A.Gi[f] = a->gi + f * a->S;
  }
  A.Hxi = a->tj;
  A.Hxix = a->tj + RJ;
  A.Het1 = a->tj + 2 * RJ;
  A.Het2 = a->tj + 3 * RJ;
  A.Hmb1 = a->tj + 4 * RJ;
  A.Hmb2 = a->tj + 5 * RJ;
  for (int k = 0; k < 14; ++k) A.O[k] = a->tf + k * RJ;
  A.Ocb1 = a->tf + 14 * RJ;
  A.Ocb2 = a->tf + 15 * RJ;
  A.part = a->part + (size_t)slot * chns::ADJ_NQ * a->mb;
  A.mb = a->mb;
  A.tw = B.tw;
  A.tabw1 = B.tabw1;
  A.tabc1 = B.tabc1;
  A.tabw2 = B.tabw2;
  A.tabc2 = B.tabc2;
  A.amask = B.amask;
  r.P = c->kp;
  return r;
}

chns::Bufs bufs_at(chns_ctx* c, chns::cd* const* in, chns::cd* const* out) {
  chns::Bufs b = c->b;
  b.W = in[0];
  b.C1 = in[1];
  b.C2 = in[2];
  b.Wo = out ? out[0] : in[0];
  b.C1o = out ? out[1] : in[1];
  b.C2o = out ? out[2] : in[2];
  return b;
}

// Forward intermediates of slot k (store mode): Hc1, Hc2, Hu, Hv, Hwx, Hwy,
// Hmu1, Hmu2, Hdx1, Hdx2 ([N][ldJ] each), then Mh1, Mh2 ([N][ldF] each).
size_t slot_entries(const chns_ctx* c) { return (size_t)c->d.N * (10 * (size_t)c->ldJ + 2 * (size_t)c->ldF); }
void point_slot(chns_ctx* c, chns::Bufs& b, int k) {
  const size_t RJ = (size_t)c->d.N * c->ldJ, RF = (size_t)c->d.N * c->ldF;
  chns::cd* q = c->adj->inter + (size_t)k * slot_entries(c);
  chns::cd** rj[] = {&b.Hc1, &b.Hc2, &b.Hu, &b.Hv, &b.Hwx, &b.Hwy, &b.Hmu1, &b.Hmu2, &b.Hdx1, &b.Hdx2};
  for (int f = 0; f < 10; ++f) *rj[f] = q + f * RJ;
  b.Mh1 = q + 10 * RJ;
  b.Mh2 = q + 10 * RJ + RF;
}

// J_N(x)^T applied to the input left in Gi/Hxi/... by the previous adj_colE,
// finished by adj_colE<stage>.  Partials go to slot (2 - stage).
// slot >= 0: passes A-C at x were kept in that intermediate slot by the
// replay; otherwise they are recomputed here into the context's buffers.
int vjp(chns_ctx* c, chns::cd* const* x, int stage, int slot) {
  const int L = c->d.log2n, M = CHNS_MODEL_TERNARY;
  hipStream_t s = c->stream;
  const chns::cd* xc[3] = {x[0], x[1], x[2]};
  chns::AdjArgs A = adj_args(c, xc, 2 - stage);
  if (slot >= 0) {
    chns::Bufs b = c->b;
    point_slot(c, b, slot);
    A.A.Hc1 = b.Hc1; A.A.Hc2 = b.Hc2; A.A.Hu = b.Hu; A.A.Hv = b.Hv; A.A.Hwx = b.Hwx; A.A.Hwy = b.Hwy;
    A.A.Hmu1 = b.Hmu1; A.A.Hmu2 = b.Hmu2; A.A.Hdx1 = b.Hdx1; A.A.Hdx2 = b.Hdx2;
    A.A.Mh1 = b.Mh1;
    A.A.Mh2 = b.Mh2;
  } else {
    const chns::Bufs b = bufs_at(c, x, nullptr);
    ALAUNCH(c, chns::launch_colA(L, b, c->kp, 2, s));
    ALAUNCH(c, chns::launch_rowB(L, M, b, c->kp, s));
    ALAUNCH(c, chns::launch_colC(L, M, b, c->kp, false, s));
  }
  ALAUNCH(c, chns::launch_adj_rowD(L, A, s));
  ALAUNCH(c, chns::launch_adj_colBC(L, A, s));
  ALAUNCH(c, chns::launch_adj_rowB(L, A, s));
  ALAUNCH(c, chns::launch_adj_colE(L, stage, A, s));
  return 0;
}

// One forward ETD2RK stage of the replay: state in -> out (pass E out of
// place).  slot >= 0: passes A-C write intermediate slot `slot` and pass E's
// fused pass A writes slot next_slot (the next stage's input); full = false
// stops after pass C (the intermediates of a stage whose output is unused).
// Passes C and D take the vorticity advection in its gradient form here
// (pass C writes dw/dx and dw/dy, which the reverse pass D reads); the
// forward sweep's conservative form is the same function of the state, so
// the replayed trajectory differs from it by round-off only.
int replay_stage(chns_ctx* c, chns::cd* const* in, chns::cd* const* out, int stage, bool need_colA, int slot,
                 int next_slot, bool full) {
  const int L = c->d.log2n, M = CHNS_MODEL_TERNARY;
  hipStream_t s = c->stream;
  chns::Bufs b = bufs_at(c, in, out);
  if (slot >= 0) point_slot(c, b, slot);
  if (need_colA) ALAUNCH(c, chns::launch_colA(L, b, c->kp, 2, s));
  ALAUNCH(c, chns::launch_rowB(L, M, b, c->kp, s));
  ALAUNCH(c, chns::launch_colC(L, M, b, c->kp, false, s));
  if (!full) return 0;
  ALAUNCH(c, chns::launch_rowD(L, M, b, c->kp, false, s));
  if (next_slot >= 0) {
    chns::Bufs nb = c->b;
    point_slot(c, nb, next_slot);
    b.Hc1 = nb.Hc1;
    b.Hc2 = nb.Hc2;
  }
  ALAUNCH(c, chns::launch_colE(L, M, stage, b, c->kp, 3, s));
  return 0;
}

int check_ctx(chns_ctx* c) {
  if (!c) return CHNS_E_ARG;
  if (c->p.model != CHNS_MODEL_TERNARY) return chns_fail(c, CHNS_E_ARG, "the adjoint supports the ternary model");
  if (c->in_ensemble) return chns_fail(c, CHNS_E_ARG, "the adjoint runs on a stand-alone context");
  if (c->p.gravity || c->p.vpm || c->p.sponge || c->p.smooth_filter || c->p.dealias != 0 || c->p.f_amp != 0.0)
    return chns_fail(c, CHNS_E_ARG,
                     "the adjoint covers the plain ternary step (no gravity, vpm, sponge, smooth_filter, dealias rule "
                     "other than 0, forcing)");
  AHIP(c, hipSetDevice(c->dev));
  if (!c->adj) c->adj = new chns_adj_state();
  return ensure_work(c);
}

// dL/dsigma from the coefficient partials (gamma_i, del of main.c:166-170;
// fac = 2 del/eps, mc_i = 0.75 g_i eps, A_ij and lambda/g_i of the forward step)
void sigma_gradient(const chns_ctx* c, const double* q, double* out) {
  const double g1 = c->d.g1, g2 = c->d.g2, g3 = c->d.g3, eps = c->d.eps, lam = c->p.lambda;
  const double D = g1 * g2 + g1 * g3 + g2 * g3, P3 = g1 * g2 * g3;
  const double dd1 = 6.0 * g2 * g3 / D - 6.0 * P3 * (g2 + g3) / (D * D);
  const double dd2 = 6.0 * g1 * g3 / D - 6.0 * P3 * (g1 + g3) / (D * D);
  const double dd3 = 6.0 * g1 * g2 / D - 6.0 * P3 * (g1 + g2) / (D * D);
  using namespace chns;
  const double rf = q[AQ_FAC] * 2.0 / eps;
  const double dg1 = q[AQ_G1] - q[AQ_G1I] / (g1 * g1) + 0.75 * eps * q[AQ_MC1] -
                     g3 / (g1 * g1) * (q[AQ_A11] + q[AQ_A21]) - lam / (g1 * g1) * q[AQ_LG1] + rf * dd1 +
                     c->adj->dgam[0];
  const double dg2 = q[AQ_G2] - q[AQ_G2I] / (g2 * g2) + 0.75 * eps * q[AQ_MC2] -
                     g3 / (g2 * g2) * (q[AQ_A12] + q[AQ_A22]) - lam / (g2 * g2) * q[AQ_LG2] + rf * dd2 +
                     c->adj->dgam[1];
  const double dg3 = q[AQ_G3] - q[AQ_G3I] / (g3 * g3) + (q[AQ_A11] + q[AQ_A21]) / g1 +
                     (q[AQ_A12] + q[AQ_A22]) / g2 + rf * dd3 + c->adj->dgam[2];
  // gamma_1 = s12 - s23 + s13, gamma_2 = s12 - s13 + s23, gamma_3 = s23 - s12 + s13
  out[0] = dg1 + dg2 - dg3;    // sigma_12
  out[1] = dg1 - dg2 + dg3;    // sigma_13
  out[2] = -dg1 + dg2 + dg3;   // sigma_23
}

// after a seed in Ub: zero the partial sums and prepare the first product's input
int after_seed(chns_ctx* c) {
  chns_adj_state* a = c->adj;
  AHIP(c, hipMemsetAsync(a->acc, 0, chns::ADJ_NQ * sizeof(double), c->stream));
  const chns::AdjArgs A = adj_args(c, nullptr, 0);
  ALAUNCH(c, chns::launch_adj_colE(c->d.log2n, 0, A, c->stream));
  a->seeded = true;
  a->backward_done = false;
  return 0;
}

}  // namespace

void chns_adj_free(chns_ctx* c) {
  if (!c || !c->adj) return;
  chns_adj_state* a = c->adj;
  (void)hipSetDevice(c->dev);
  (void)hipStreamSynchronize(c->stream);
  afree(a->ckpt);
  afree(a->traj);
  afree(a->inter);
  afree(a->ub);
  afree(a->ab);
  afree(a->gi);
  afree(a->tj);
  afree(a->tf);
  afree(a->part);
  afree(a->acc);
  afree(a->dS);
  afree(a->nbq);
  delete a;
  c->adj = nullptr;
}

extern "C" int chns_adj_set_store(chns_ctx* c, int store) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->adj->store = store ? 1 : 0;
  if (!store) {
    afree(c->adj->inter);
    c->adj->inter_K = 0;
  }
  return CHNS_OK;
}

extern "C" int chns_adj_release(chns_ctx* c) {
  if (!c) return CHNS_E_ARG;
  chns_adj_free(c);
  return CHNS_OK;
}

extern "C" int chns_adj_forward(chns_ctx* c, int nsteps, int ckpt_every) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_state) return chns_fail(c, CHNS_E_STATE, "no state set");
  if (nsteps < 0 || ckpt_every < 1) return chns_fail(c, CHNS_E_ARG, "nsteps >= 0 and ckpt_every >= 1 required");
  chns_adj_state* a = c->adj;
  const int K = nsteps < ckpt_every ? (nsteps > 0 ? nsteps : 1) : ckpt_every;
  const int nseg = nsteps > 0 ? (nsteps + K - 1) / K : 0;
  if (a->ckpt_n < nseg) {
    afree(a->ckpt);
    a->ckpt_n = 0;
    if ((rc = amalloc(c, &a->ckpt, (size_t)nseg * 3 * a->S))) return rc;
    a->ckpt_n = nseg;
  }
  if (a->traj_K < K) {
    afree(a->traj);
    a->traj_K = 0;
    if ((rc = amalloc(c, &a->traj, (size_t)(2 * K + 1) * 3 * a->S))) return rc;
    a->traj_K = K;
  }
  // intermediates of a segment: kept when they fit in 85 % of the free memory
  if (a->store && a->inter_K < K) {
    afree(a->inter);
    a->inter_K = 0;
    size_t fr = 0, tot = 0;
    const size_t need = (size_t)2 * K * slot_entries(c) * sizeof(chns::cd);
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && need < fr / 100 * 85) {
      void* q = nullptr;
      if (hipMalloc(&q, need) == hipSuccess) {
        a->inter = (chns::cd*)q;
        a->inter_K = K;
      } else {
        (void)hipGetLastError();   // fall back to recomputing passes A-C
      }
    }
  }
  a->nsteps = nsteps;
  a->K = K;
  a->nseg = nseg;
  chns::cd* st[3] = {c->b.W, c->b.C1, c->b.C2};
  for (int sg = 0; sg < nseg; ++sg) {
    for (int f = 0; f < 3; ++f)
      AHIP(c, hipMemcpyAsync(field(a->ckpt, a->S, sg, f), st[f], a->S * sizeof(chns::cd), hipMemcpyDeviceToDevice,
                             c->stream));
    const int len = (sg + 1) * K <= nsteps ? K : nsteps - sg * K;
    if ((rc = chns_step(c, len))) return rc;
  }
  a->forward_done = true;
  a->seeded = false;
  a->backward_done = false;
  return CHNS_OK;
}

extern "C" int chns_adj_seed_spectrum(chns_ctx* c, const double* target, double* loss, double* sph1) {
  int rc = check_ctx(c);
  if (rc) return rc;
  chns_adj_state* a = c->adj;
  if (!a->forward_done) return chns_fail(c, CHNS_E_STATE, "chns_adj_forward first");
  if (!target) return chns_fail(c, CHNS_E_ARG, "null target spectrum");
  std::vector<double> S(a->nsh), d(a->nsh);
  if ((rc = chns_spectrum(c, 0, nullptr, nullptr, S.data(), nullptr, nullptr, nullptr, nullptr))) return rc;
  double L = 0.0;
  for (int k = 0; k < a->nsh; ++k) {
    d[k] = S[k] - target[k];
    L += d[k] * d[k];
  }
  if (loss) *loss = L;
  if (sph1) std::memcpy(sph1, S.data(), a->nsh * sizeof(double));
  for (double& x : a->dgam) x = 0.0;   // the spectrum does not depend on sigma directly
  AHIP(c, hipMemcpyAsync(a->dS, d.data(), a->nsh * sizeof(double), hipMemcpyHostToDevice, c->stream));
  chns::cd* ubp[3] = {a->ub, a->ub + a->S, a->ub + 2 * a->S};
  ALAUNCH(c, chns::launch_adj_seed_spectrum(c->b.C1, ubp, a->dS, c->b.amask, c->d.N, c->d.factor_x,
                                            2.0 * M_PI / c->p.lx, a->nsh, c->stream));
  if ((rc = after_seed(c))) return rc;
  AHIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

extern "C" int chns_adj_seed_free_energy(chns_ctx* c, double* loss, double* ephi, double* egphi) {
  int rc = check_ctx(c);
  if (rc) return rc;
  chns_adj_state* a = c->adj;
  if (!a->forward_done) return chns_fail(c, CHNS_E_STATE, "chns_adj_forward first");
  if ((rc = chns_diag_prepare(c))) return rc;
  const int N = c->d.N, L = c->d.log2n;
  const double g1 = c->d.g1, g2 = c->d.g2, g3 = c->d.g3, eps = c->d.eps;
  const double scale = 1.0 / ((double)c->p.Nx * c->p.Ny);
  // T(c1), T(c2) and their gradients in real space (every read of the state
  // inside the step is truncated; so is this one)
  double* f[6];
  if ((rc = chns_fe_fields(c, true, f))) return rc;
  double q[chns::RED_Q];
  if ((rc = chns_fe_sums(c, f, g1, g2, g3, q))) return rc;
  const double e_bulk = q[0] * (6.0 / eps) * scale;     // free_energy.c:246-248
  const double e_grad = q[1] * (0.375 * eps * scale);   // free_energy.c:335
  // gamma_i weights a sum of its own in both energies: dL/dgamma_i directly
  for (int k = 0; k < 3; ++k) {
    if ((rc = chns_fe_sums(c, f, k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0, q))) return rc;
    a->dgam[k] = q[0] * (6.0 / eps) * scale + q[1] * (0.375 * eps * scale);
  }
  // real-space adjoints in place, their r2c (truncated to the kept modes),
  // then c_i bar = X_c - i kx X_x - i ky X_y
  ALAUNCH(c, chns::launch_fe_adjoint(f, N, g1, g2, g3, (6.0 / eps) * scale, 0.375 * eps * scale, c->stream));
  chns::cd* X[6] = {a->ab, a->ab + a->S, a->ab + 2 * a->S, a->gi, a->gi + a->S, a->gi + 2 * a->S};
  for (int k = 0; k < 3; ++k) {
    ALAUNCH(c, chns::launch_row_r2c(L, f[2 * k], f[2 * k + 1], c->dT1, c->dT2, c->b.tw, c->ldF, c->stream));
    ALAUNCH(c, chns::launch_col_fwd(L, true, c->dT1, X[2 * k], c->b.tw, c->b.amask, nullptr, c->ldF, c->stream));
    ALAUNCH(c, chns::launch_col_fwd(L, true, c->dT2, X[2 * k + 1], c->b.tw, c->b.amask, nullptr, c->ldF, c->stream));
  }
  chns::cd* ubp[3] = {a->ub, a->ub + a->S, a->ub + 2 * a->S};
  ALAUNCH(c, chns::launch_fe_seed(X, ubp, N, c->d.factor_x, c->stream));
  if ((rc = after_seed(c))) return rc;
  AHIP(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = e_bulk + e_grad;
  if (ephi) *ephi = e_bulk;
  if (egphi) *egphi = e_grad;
  return CHNS_OK;
}

extern "C" int chns_adj_seed(chns_ctx* c, const double* wbar, const double* c1bar, const double* c2bar) {
  int rc = check_ctx(c);
  if (rc) return rc;
  chns_adj_state* a = c->adj;
  if (!a->forward_done) return chns_fail(c, CHNS_E_STATE, "chns_adj_forward first");
  const int N = c->d.N, H = c->d.H;
  const double* src[3] = {wbar, c1bar, c2bar};
  for (double& x : a->dgam) x = 0.0;   // a caller-given seed has no direct sigma dependence
  if ((rc = chns_stage_buffer(c))) return rc;
  for (int f = 0; f < 3; ++f) {
    chns::cd* dst = a->ub + f * a->S;
    if (src[f]) {
      // IDXC in, transposed to the column-major state on the device
      AHIP(c, hipMemcpyAsync(c->dstage, src[f], a->S * sizeof(chns::cd), hipMemcpyHostToDevice, c->stream));
      ALAUNCH(c, chns::launch_transpose(c->dstage, dst, N, H, c->stream));
      ALAUNCH(c, chns::launch_truncate(dst, c->b.amask, nullptr, N, c->stream));   // adjoint states live on the kept modes
    } else {
      AHIP(c, hipMemsetAsync(dst, 0, a->S * sizeof(chns::cd), c->stream));
    }
  }
  if ((rc = after_seed(c))) return rc;
  AHIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

extern "C" int chns_adj_backward(chns_ctx* c, double* g_wz, double* g_phi1, double* g_phi2, double* g_sigma) {
  int rc = check_ctx(c);
  if (rc) return rc;
  chns_adj_state* a = c->adj;
  if (!a->seeded) return chns_fail(c, CHNS_E_STATE, "seed the adjoint first");
  const int K = a->K;
  hipStream_t s = c->stream;
  for (int sg = a->nseg - 1; sg >= 0; --sg) {
    const int len = (sg + 1) * K <= a->nsteps ? K : a->nsteps - sg * K;
    // replay the segment from its checkpoint, keeping u_k and a_k
    chns::cd* u[3];
    chns::cd* av[3];
    chns::cd* un[3];
    const bool st = a->inter && a->inter_K >= K;
    for (int k = 0; k < len; ++k) {
      for (int f = 0; f < 3; ++f) {
        u[f] = k == 0 ? field(a->ckpt, a->S, sg, f) : field(a->traj, a->S, k, f);
        av[f] = field(a->traj, a->S, K + 1 + k, f);
        un[f] = field(a->traj, a->S, k + 1, f);
      }
      const bool last = k + 1 == len;
      if ((rc = replay_stage(c, u, av, 1, k == 0, st ? 2 * k : -1, st ? 2 * k + 1 : -1, true))) return rc;
      // the last step's stage 2 only matters for its intermediates (store mode)
      if ((!last || st) &&
          (rc = replay_stage(c, av, un, 2, false, st ? 2 * k + 1 : -1, st && !last ? 2 * k + 2 : -1, !last)))
        return rc;
    }
    // reverse sweep over the segment
    for (int k = len - 1; k >= 0; --k) {
      for (int f = 0; f < 3; ++f) {
        u[f] = k == 0 ? field(a->ckpt, a->S, sg, f) : field(a->traj, a->S, k, f);
        av[f] = field(a->traj, a->S, K + 1 + k, f);
      }
      if ((rc = vjp(c, av, 2, st ? 2 * k + 1 : -1)) || (rc = vjp(c, u, 1, st ? 2 * k : -1))) return rc;
      ALAUNCH(c, chns::launch_adj_accum(a->part, 2, (size_t)chns::ADJ_NQ * a->mb, a->mb, a->nbq, a->acc, s));
    }
  }
  c->hc_valid = false;   // pass A's outputs no longer belong to the current state
  a->backward_done = true;
  a->seeded = false;
  double q[chns::ADJ_NQ];
  AHIP(c, hipMemcpyAsync(q, a->acc, sizeof(q), hipMemcpyDeviceToHost, s));
  AHIP(c, hipStreamSynchronize(s));
  if (g_sigma) sigma_gradient(c, q, g_sigma);
  double* outs[3] = {g_wz, g_phi1, g_phi2};
  for (int f = 0; f < 3; ++f)
    if (outs[f] && (rc = chns_real_of(c, a->ub + f * a->S, outs[f]))) return rc;
  return CHNS_OK;
}

extern "C" int chns_adj_get(chns_ctx* c, double* wbar, double* c1bar, double* c2bar, double* partials) {
  int rc = check_ctx(c);
  if (rc) return rc;
  chns_adj_state* a = c->adj;
  if (!a->backward_done) return chns_fail(c, CHNS_E_STATE, "no completed backward sweep");
  const int N = c->d.N, H = c->d.H;
  double* dst[3] = {wbar, c1bar, c2bar};
  if ((rc = chns_stage_buffer(c))) return rc;
  for (int f = 0; f < 3; ++f) {
    if (!dst[f]) continue;
    ALAUNCH(c, chns::launch_transpose(a->ub + f * a->S, c->dstage, H, N, c->stream));
    AHIP(c, hipMemcpyAsync(dst[f], c->dstage, a->S * sizeof(chns::cd), hipMemcpyDeviceToHost, c->stream));
  }
  AHIP(c, hipStreamSynchronize(c->stream));
  if (partials) {
    AHIP(c, hipMemcpyAsync(partials, a->acc, chns::ADJ_NQ * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    AHIP(c, hipStreamSynchronize(c->stream));
  }
  return CHNS_OK;
}

// C ABI of the CHNS-2D ETD2RK hot path: context, device buffers, state
// transfer and the per-step launch sequence.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "chns.h"
#include "chns_host.hpp"
#include "chns_launch.hpp"
#include "chns_types.hpp"

#include "chns_ctx.hpp"

int chns_fail(chns_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int chns_hip_fail(chns_ctx* c, hipError_t e, const char* what) {
  return chns_fail(c, CHNS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

bool chns_graphs_default() {
  const char* v = std::getenv("CHNS_NO_GRAPH");
  return !(v && v[0] == '1');
}

int chns_dalloc_bytes(chns_ctx* c, void** p, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return chns_hip_fail(c, e, "hipMalloc");
  c->allocs.push_back(q);
  e = hipMemset(q, 0, bytes);
  if (e != hipSuccess) return chns_hip_fail(c, e, "hipMemset");
  *p = q;
  return 0;
}

namespace {

int fail(chns_ctx* c, int code, const std::string& msg) { return chns_fail(c, code, msg); }
int hip_fail(chns_ctx* c, hipError_t e, const char* what) { return chns_hip_fail(c, e, what); }

#define CHNS_HIP(ctx, call)                                          \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);         \
  } while (0)

#define CHNS_LAUNCH(ctx, call)                                                        \
  do {                                                                                \
    int e_ = (call);                                                                  \
    if (e_ != 0) return hip_fail((ctx), (hipError_t)e_, #call);                      \
  } while (0)

int round_up(int a, int b) { return (a + b - 1) / b * b; }

template <class T> int dalloc(chns_ctx* c, T** p, size_t count) {
  return chns_dalloc_bytes(c, (void**)p, count * sizeof(T));
}

}  // namespace

namespace {

void fill_kparams(chns_ctx* c) {
  const chns_params& p = c->p;
  const chns_host::Derived& d = c->d;
  chns::KParams& k = c->kp;
  k.N = d.N;
  k.ldJ = c->ldJ;
  k.ldF = c->ldF;
  k.Jt = c->Jt;
  k.f_k = p.f_amp != 0.0 ? p.f_k : 0;
  k.f = d.factor_x;
  k.scale = 1.0 / ((double)p.Nx * p.Ny);                  // fft.c:21
  k.force_im = -0.5 * p.f_amp * (double)p.Nx * (double)p.Ny;
  const double eps = d.eps;
  k.g1 = d.g1; k.g2 = d.g2; k.g3 = d.g3;
  if (p.model == CHNS_MODEL_TERNARY) {
    k.g1i = 1.0 / d.g1; k.g2i = 1.0 / d.g2; k.g3i = 1.0 / d.g3;   // chem_potential.c:21-23
    k.fac = 2.0 * d.del / eps;                                    // chem_potential.c:85-86
    k.X6 = 6.0 * p.Xlambda;                                       // chem_potential.c:78-80
    k.mc1 = 0.75 * d.g1 * eps;                                    // surface_tension.c:123
    k.mc2 = 0.75 * d.g2 * eps;                                    // surface_tension.c:124
    k.A11 = (1.0 + (d.g3 / d.g1)); k.A12 = (d.g3 / d.g2);         // surface_tension.c:144
    k.A21 = (d.g3 / d.g1); k.A22 = (1.0 + (d.g3 / d.g2));         // surface_tension.c:195
    k.lg1 = p.lambda / d.g1; k.lg2 = p.lambda / d.g2;             // nonlin.c:369-370
  }
  k.grav = p.gravity ? 1 : 0;
  k.vpm = p.vpm ? 1 : 0;
  k.sponge = p.sponge ? 1 : 0;
  if (p.model == CHNS_MODEL_TERNARY) {
    k.gr1 = (-1.0 * p.g) * (p.rho1 - p.rho3);                     // nonlin.c:151-154
    k.gr2 = (-1.0 * p.g) * (p.rho2 - p.rho3);
  } else {
    k.gr1 = (-1.0 * p.g) * (p.rho1 - p.rho2);                     // nonlin.c:117-118
    k.gr2 = 0.0;
  }
  k.etainv = 1.0 / (1.5 * p.dt);                                  // global.h:48-49
  k.mcb = 1.5 * p.sigma * eps;                                    // surface_tension.c:141
  k.s24 = 24.0 * (p.sigma / eps);                                 // surface_tension.c:142, nonlin.c:240
  k.lam = p.lambda;                                               // nonlin.c:251
}

// Launch through `fn`, bracketed by a pair of events when timing is on.
template <class F> int timed(chns_ctx* c, int kind, F fn) {
  if (!c->timing) return fn();
  while (c->ev_pool.size() < 2 * (c->ev_used + 1)) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return (int)r;
    c->ev_pool.push_back(e);
  }
  hipEvent_t a = c->ev_pool[2 * c->ev_used], b = c->ev_pool[2 * c->ev_used + 1];
  hipError_t r = hipEventRecord(a, c->stream);
  if (r != hipSuccess) return (int)r;
  int rc = fn();
  if (rc) return rc;
  r = hipEventRecord(b, c->stream);
  if (r != hipSuccess) return (int)r;
  c->ev_kind.push_back(kind);
  ++c->ev_used;
  return 0;
}

}  // namespace

extern "C" const char* chns_last_error(const chns_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" void chns_destroy(chns_ctx* c) {
  if (!c || c->in_ensemble) return;   // an ensemble's instances go with the ensemble
  // queued work (graph replays included) finishes before anything it uses goes
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  chns_adj_free(c);
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (void* q : c->allocs) (void)hipFree(q);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int chns_create(chns_ctx** out, const chns_params* p, int device) {
  return chns_create_on(out, p, device, nullptr);
}

int chns_create_on(chns_ctx** out, const chns_params* p, int device, hipStream_t shared) {
  if (!out) return CHNS_E_ARG;
  *out = nullptr;
  if (const char* why = chns_host::check_params(p)) {
    std::fprintf(stderr, "chns_create: %s\n", why);
    return CHNS_E_ARG;
  }
  chns_ctx* c = new chns_ctx();
  c->p = *p;
  c->dev = device;
  c->graphs = chns_graphs_default();
  chns_host::derive(p, &c->d);
  if (c->p.epsilon <= 0.0) c->p.epsilon = c->d.eps;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    std::fprintf(stderr, "chns_create: hipSetDevice(%d): %s\n", device, hipGetErrorString(e));
    delete c;
    return CHNS_E_HIP;
  }
  int rc = 0;
  auto bail = [&](int r) { int code = r; std::fprintf(stderr, "chns_create: %s\n", c->err.c_str()); chns_destroy(c); return code; };
  if ((rc = chns_noalloc_setup(device, nullptr)) != 0)
    return bail(hip_fail(c, (hipError_t)rc, "hipGetSymbolAddress(scratch)"));
  if (shared) {
    c->stream = shared;
    c->own_stream = false;
  } else {
    e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipStreamCreate"));
  }

  const int N = c->d.N, H = c->d.H;
  c->amask.assign(H, -1);
  chns_host::dealias_extent(p, c->amask.data());
  c->Jt = 0;
  for (int j = 0; j < H; ++j)
    if (c->amask[j] >= 0) c->Jt = j + 1;
  if (c->Jt == 0) return bail(fail(c, CHNS_E_ARG, "dealias mask keeps no mode"));
  c->ldJ = round_up(c->Jt, 4);
  c->ldF = round_up(H, 4);
  c->nph = p->model == CHNS_MODEL_TERNARY ? 2 : (p->model == CHNS_MODEL_BINARY ? 1 : 0);
  fill_kparams(c);

  chns::Bufs& b = c->b;
  const size_t S = (size_t)N * H, RJ = (size_t)N * c->ldJ, RF = (size_t)N * c->ldF;
  if ((rc = dalloc(c, &b.W, S)) || (rc = dalloc(c, &b.NW, S))) return bail(rc);
  if (c->nph >= 1 && ((rc = dalloc(c, &b.C1, S)) || (rc = dalloc(c, &b.NC1, S)))) return bail(rc);
  if (c->nph >= 2 && ((rc = dalloc(c, &b.C2, S)) || (rc = dalloc(c, &b.NC2, S)))) return bail(rc);
  b.Wo = b.W;
  b.C1o = b.C1;
  b.C2o = b.C2;
  chns::cd** rj[] = {&b.Hc1, &b.Hc2, &b.Hu, &b.Hv, &b.Hwx, &b.Hwy, &b.Hmu1, &b.Hmu2, &b.Hdx1, &b.Hdx2};
  for (auto q : rj)
    if ((rc = dalloc(c, q, RJ))) return bail(rc);
  chns::cd** rf[] = {&b.Mh1, &b.Mh2, &b.G1, &b.G2, &b.G3, &b.G4};
  for (auto q : rf)
    if ((rc = dalloc(c, q, RF))) return bail(rc);

  const size_t HH = (size_t)H * H;
  std::vector<double> tw(2 * (size_t)N), tw1(2 * HH), tc1(2 * HH), tw2(HH), tc2(HH);
  chns_host::build_twiddles(N, tw.data());
  chns_host::build_coef_tables(&c->p, tw1.data(), tc1.data(), tw2.data(), tc2.data());
  chns::cd *dtw, *dw1, *dc1;
  double *dw2, *dc2;
  int* dam;
  if ((rc = dalloc(c, &dtw, N)) || (rc = dalloc(c, &dw1, HH)) || (rc = dalloc(c, &dc1, HH)) ||
      (rc = dalloc(c, &dw2, HH)) || (rc = dalloc(c, &dc2, HH)) || (rc = dalloc(c, &dam, H)))
    return bail(rc);
  if ((e = hipMemcpy(dtw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dw1, tw1.data(), tw1.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dc1, tc1.data(), tc1.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dw2, tw2.data(), tw2.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dc2, tc2.data(), tc2.size() * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dam, c->amask.data(), H * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return bail(hip_fail(c, e, "hipMemcpy(tables)"));
  if (p->vpm) {
    std::vector<double> chi(N), spg(N);
    chns_host::build_wall_tables(p, chi.data(), spg.data());
    double *dchi, *dspg;
    if ((rc = dalloc(c, &dchi, N)) || (rc = dalloc(c, &dspg, N))) return bail(rc);
    if ((e = hipMemcpy(dchi, chi.data(), N * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dspg, spg.data(), N * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(hip_fail(c, e, "hipMemcpy(walls)"));
    b.chi = dchi;
    b.spg = dspg;
  }
  if (p->smooth_filter || p->dealias != 0) {
    // the smooth filter's factors, or 1/0 on the kept/cut modes of the 2/3
    // rule and the N/3 circle: both run the unpruned (table) kernels
    std::vector<double> ft(HH);
    if (p->smooth_filter) chns_host::build_filter_table(p, ft.data());
    else chns_host::build_cut_table(p, c->amask.data(), ft.data());
    double* dft;
    if ((rc = dalloc(c, &dft, HH))) return bail(rc);
    if ((e = hipMemcpy(dft, ft.data(), HH * 8, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(hip_fail(c, e, "hipMemcpy(filter)"));
    b.filt = dft;
  }
  if ((rc = dalloc(c, &b.flag, 1))) return bail(rc);
  if (p->gravity && (rc = dalloc(c, &b.GR, S))) return bail(rc);
  // T(w) for the sponge, and for pass D's conservative vorticity advection
  if ((p->sponge || c->nph > 0) && (rc = dalloc(c, &b.Hw, RJ))) return bail(rc);
  b.tw = dtw;
  b.tabw1 = dw1;
  b.tabc1 = dc1;
  b.tabw2 = dw2;
  b.tabc2 = dc2;
  b.amask = dam;
  *out = c;
  return CHNS_OK;
}

// Device buffer for one field in the reference layout, allocated on first use.
int chns_stage_buffer(chns_ctx* c) {
  if (c->dstage) return CHNS_OK;
  return chns_dalloc_bytes(c, (void**)&c->dstage, (size_t)c->d.N * c->d.H * sizeof(chns::cd));
}

// The state crosses the boundary in the reference layout IDXC [Nx][Ny/2+1]
// (global.h:13): each field is copied as it is into a device staging buffer
// (or out of it) and transposed to or from the column-major [Ny/2+1][Nx]
// layout on the device.
extern "C" int chns_set_state(chns_ctx* c, const double* cwz, const double* cphi1, const double* cphi2) {
  if (!c) return CHNS_E_ARG;
  if (!c->have_state && (!cwz || (c->nph >= 1 && !cphi1) || (c->nph >= 2 && !cphi2)))
    return fail(c, CHNS_E_ARG, "null state array (the first upload needs every field)");
  CHNS_HIP(c, hipSetDevice(c->dev));
  const int rc = chns_stage_buffer(c);
  if (rc != CHNS_OK) return rc;
  const int N = c->d.N, H = c->d.H;
  const size_t bytes = (size_t)N * H * sizeof(chns::cd);
  const double* src[3] = {cwz, cphi1, cphi2};
  chns::cd* dst[3] = {c->b.W, c->b.C1, c->b.C2};
  for (int f = 0; f <= c->nph; ++f) {
    if (!src[f]) continue;   // keeps the field on the device
    CHNS_HIP(c, hipMemcpyAsync(c->dstage, src[f], bytes, hipMemcpyHostToDevice, c->stream));
    CHNS_LAUNCH(c, chns::launch_transpose(c->dstage, dst[f], N, H, c->stream));
  }
  CHNS_HIP(c, hipMemsetAsync(c->b.flag, 0, sizeof(int), c->stream));
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  c->have_state = true;
  c->hc_valid = false;
  c->energy_flag = 0;
  return CHNS_OK;
}

extern "C" int chns_get_state(chns_ctx* c, double* cwz, double* cphi1, double* cphi2) {
  if (!c) return CHNS_E_ARG;
  if (!c->have_state) return fail(c, CHNS_E_STATE, "no state set");
  CHNS_HIP(c, hipSetDevice(c->dev));
  const int rc = chns_stage_buffer(c);
  if (rc != CHNS_OK) return rc;
  const int N = c->d.N, H = c->d.H;
  const size_t bytes = (size_t)N * H * sizeof(chns::cd);
  double* dsts[3] = {cwz, cphi1, cphi2};
  const chns::cd* src[3] = {c->b.W, c->b.C1, c->b.C2};
  for (int f = 0; f <= c->nph; ++f) {
    if (!dsts[f]) continue;
    CHNS_LAUNCH(c, chns::launch_transpose(src[f], c->dstage, H, N, c->stream));
    CHNS_HIP(c, hipMemcpyAsync(dsts[f], c->dstage, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

extern "C" int chns_set_real_state(chns_ctx* c, const double* wz, const double* phi1, const double* phi2) {
  if (!c) return CHNS_E_ARG;
  if ((c->nph >= 1 && !phi1) || (c->nph >= 2 && !phi2)) return fail(c, CHNS_E_ARG, "null phase field");
  CHNS_HIP(c, hipSetDevice(c->dev));
  const int N = c->d.N;
  const size_t nb = (size_t)N * N * sizeof(double);
  // The row-pass output buffers are free between steps: use Mh1/Mh2 as real
  // staging ([N][N] doubles fit in [N][ldF] complex) and G1/G2 as the
  // half-transformed rows.
  double* ra = (double*)c->b.Mh1;
  double* rb = (double*)c->b.Mh2;
  auto dfft_trunc = [&](const double* a, const double* bb, chns::cd* oa, chns::cd* ob) -> int {
    CHNS_HIP(c, hipMemcpyAsync(ra, a, nb, hipMemcpyHostToDevice, c->stream));
    if (bb) CHNS_HIP(c, hipMemcpyAsync(rb, bb, nb, hipMemcpyHostToDevice, c->stream));
    CHNS_LAUNCH(c, chns::launch_row_r2c(c->d.log2n, ra, bb ? rb : nullptr, c->b.G1, c->b.G2, c->b.tw, c->ldF,
                                        c->stream));
    CHNS_LAUNCH(c, chns::launch_col_fwd(c->d.log2n, true, c->b.G1, oa, c->b.tw, c->b.amask, c->b.filt, c->ldF, c->stream));
    if (bb) CHNS_LAUNCH(c, chns::launch_col_fwd(c->d.log2n, true, c->b.G2, ob, c->b.tw, c->b.amask, c->b.filt, c->ldF,
                                                c->stream));
    CHNS_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
  };
  int rc;
  if (wz) {
    if ((rc = dfft_trunc(wz, nullptr, c->b.W, nullptr))) return rc;
  } else {
    CHNS_HIP(c, hipMemsetAsync(c->b.W, 0, (size_t)N * c->d.H * sizeof(chns::cd), c->stream));
    CHNS_HIP(c, hipStreamSynchronize(c->stream));
  }
  if (c->nph == 1 && (rc = dfft_trunc(phi1, nullptr, c->b.C1, nullptr))) return rc;
  if (c->nph == 2 && (rc = dfft_trunc(phi1, phi2, c->b.C1, c->b.C2))) return rc;
  CHNS_HIP(c, hipMemsetAsync(c->b.flag, 0, sizeof(int), c->stream));
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  c->have_state = true;
  c->hc_valid = false;
  c->energy_flag = 0;
  return CHNS_OK;
}

namespace {

// The launches of one ETD2RK step (etd2rk.c:14-224): two stages of passes
// B, C, D (and the buoyancy term) and E.  Each pass is bracketed by timing
// events when timing is on.
int enqueue_step(chns_ctx* c) {
  const int L = c->d.log2n, M = c->p.model;
  hipStream_t s = c->stream;
  const chns::Bufs& B = c->b;
  const chns::KParams& K = c->kp;
  // SMOOTH_FILTER: the passes of chns_filter.hip
  const bool F = B.filt != nullptr;
  // passes C and D take the vorticity advection in conservative form, except
  // with the sponge, whose term has no derivative to share
  const bool cons = !K.sponge;
  for (int st = 1; st <= 2; ++st) {
    if (c->nph > 0)
      CHNS_LAUNCH(c, timed(c, CHNS_PASS_ROWB, [&] {
        return F ? chns::launch_filt_rowB(L, M, B, K, s) : chns::launch_rowB(L, M, B, K, s);
      }));
    CHNS_LAUNCH(c, timed(c, CHNS_PASS_COLC, [&] {
      return F ? chns::launch_filt_colC(L, M, B, K, s) : chns::launch_colC(L, M, B, K, cons, s);
    }));
    CHNS_LAUNCH(c, timed(c, CHNS_PASS_ROWD, [&] {
      return F ? chns::launch_filt_rowD(L, M, B, K, s) : chns::launch_rowD(L, M, B, K, cons, s);
    }));
    if (K.grav) CHNS_LAUNCH(c, chns::launch_gravity(B.C1, c->nph == 2 ? B.C2 : nullptr, B.GR, K.gr1, K.gr2, K.f,
                                                    c->d.N, s));
    CHNS_LAUNCH(c, timed(c, st == 1 ? CHNS_PASS_COLE1 : CHNS_PASS_COLE2, [&] {
      return F ? chns::launch_filt_colE(L, M, st, B, K, 1 + c->nph, s)
               : chns::launch_colE(L, M, st, B, K, 1 + c->nph, s);
    }));
  }
  return CHNS_OK;
}

// Steps per replay of the step graph.  Measured on MI355X against direct
// launches: a graph of one step is 2-5 us/step slower at 128^2-1024^2, a
// graph of 8 steps 3 % faster at 128^2 and equal from 512^2 up.
constexpr int kGraphSteps = 8;

// Capture kGraphSteps steps into c->step_exec (kernel nodes only: no memset
// or copy) and upload it, so its first replay starts like the later ones.
int capture_step(chns_ctx* c) {
  CHNS_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = CHNS_OK;
  for (int k = 0; k < kGraphSteps && rc == CHNS_OK; ++k) rc = enqueue_step(c);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc != CHNS_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != CHNS_OK ? rc : hip_fail(c, e, "hipStreamEndCapture");
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(c, ei, "hipGraphInstantiate");
  }
  c->step_graph = graph;
  c->step_exec = exec;
  CHNS_HIP(c, hipGraphUpload(exec, c->stream));
  return CHNS_OK;
}

}  // namespace

// Steps are replayed kGraphSteps at a time from a graph of captured steps,
// the remainder launched directly; the graph is captured on the first call
// whatever its length, so a short first call (a warm-up) leaves it ready.
// With event timing on, graphs off, or when the caller is capturing the
// stream into a graph of its own, every pass is launched directly.
extern "C" int chns_step(chns_ctx* c, int nsteps) {
  if (!c) return CHNS_E_ARG;
  if (nsteps < 0) return fail(c, CHNS_E_ARG, "nsteps < 0");
  if (!c->have_state) return fail(c, CHNS_E_STATE, "no state set");
  if (nsteps == 0) return CHNS_OK;
  CHNS_HIP(c, hipSetDevice(c->dev));
  const chns::Bufs& B = c->b;
  const bool F = B.filt != nullptr;
  if (c->nph > 0 && !c->hc_valid)
    CHNS_LAUNCH(c, timed(c, CHNS_PASS_COLA, [&] {
      return F ? chns::launch_filt_colA(c->d.log2n, B, c->kp, c->nph, c->stream)
               : chns::launch_colA(c->d.log2n, B, c->kp, c->nph, c->stream);
    }));
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CHNS_HIP(c, hipStreamIsCapturing(c->stream, &cap));
  if (c->timing || !c->graphs || cap != hipStreamCaptureStatusNone) {
    for (int n = 0; n < nsteps; ++n) {
      const int rc = enqueue_step(c);
      if (rc != CHNS_OK) return rc;
    }
  } else {
    int n = 0;
    if (!c->step_exec) {
      const int rc = capture_step(c);
      if (rc != CHNS_OK) return rc;
    }
    for (; n + kGraphSteps <= nsteps; n += kGraphSteps) CHNS_HIP(c, hipGraphLaunch(c->step_exec, c->stream));
    for (; n < nsteps; ++n) {
      const int rc = enqueue_step(c);
      if (rc != CHNS_OK) return rc;
    }
  }
  c->hc_valid = true;
  return CHNS_OK;
}

extern "C" int chns_sync(chns_ctx* c) {
  if (!c) return CHNS_E_ARG;
  CHNS_HIP(c, hipSetDevice(c->dev));
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  return CHNS_OK;
}

extern "C" int chns_truncate_state(chns_ctx* c) {
  if (!c) return CHNS_E_ARG;
  if (!c->have_state) return fail(c, CHNS_E_STATE, "no state set");
  CHNS_HIP(c, hipSetDevice(c->dev));
  if (c->nph >= 1) CHNS_LAUNCH(c, chns::launch_truncate(c->b.C1, c->b.amask, c->b.filt, c->d.N, c->stream));
  if (c->nph >= 2) CHNS_LAUNCH(c, chns::launch_truncate(c->b.C2, c->b.amask, c->b.filt, c->d.N, c->stream));
  // Pass A's outputs (the column iFFTs of T(c_i)) were made from the state
  // before this truncation: the next step rebuilds them from T(T(c_i)).  With
  // the sharp cut T is idempotent and they would not change; with the smooth
  // filter T(T(c)) multiplies by the factor twice more (spectrum.c:24-26, then
  // chem_potential.c:27-32 at the next step).
  if (c->nph >= 1) c->hc_valid = false;
  return CHNS_OK;
}

extern "C" int chns_graphs_enable(chns_ctx* c, int enable) {
  if (!c) return CHNS_E_ARG;
  c->graphs = enable != 0;
  return CHNS_OK;
}

extern "C" int chns_status(chns_ctx* c, int* flags) {
  if (!c) return CHNS_E_ARG;
  CHNS_HIP(c, hipSetDevice(c->dev));
  int dev_flag = 0;
  CHNS_HIP(c, hipMemcpyAsync(&dev_flag, c->b.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  const int f = (dev_flag ? CHNS_FLAG_NONFINITE : 0) | (c->energy_flag ? CHNS_FLAG_ENERGY : 0);
  if (flags) *flags = f;
  if (f == 0) return CHNS_OK;
  return fail(c, CHNS_E_DIVERGED, (f & CHNS_FLAG_NONFINITE) ? "a step wrote a non-finite value into the state"
                                                            : "bulk free energy ephi >= 100 (free_energy.c:151-154)");
}

extern "C" int chns_timing_enable(chns_ctx* c, int enable) {
  if (!c) return CHNS_E_ARG;
  c->timing = enable != 0;
  return CHNS_OK;
}

extern "C" int chns_timing_read(chns_ctx* c, double* total_ms, int* launches) {
  if (!c || !total_ms || !launches) return CHNS_E_ARG;
  CHNS_HIP(c, hipSetDevice(c->dev));
  CHNS_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < CHNS_NPASS; ++k) { total_ms[k] = 0.0; launches[k] = 0; }
  for (size_t e = 0; e < c->ev_used; ++e) {
    float ms = 0.f;
    CHNS_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[2 * e], c->ev_pool[2 * e + 1]));
    total_ms[c->ev_kind[e]] += ms;
    launches[c->ev_kind[e]] += 1;
  }
  c->ev_used = 0;
  c->ev_kind.clear();
  return CHNS_OK;
}

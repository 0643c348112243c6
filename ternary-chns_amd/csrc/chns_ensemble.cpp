// Ensembles: many independent CHNS-2D instances of mixed sizes and models on
// one GPU.  Instances of equal (N, model) form a group whose passes run as one
// batched launch each (grid.z = instance); groups run concurrently on their
// own streams, and one ETD2RK step of the whole ensemble is captured once
// into a HIP graph that is replayed per step.  Each instance is a full
// chns_ctx on the ensemble's stream, so state transfer, diagnostics and
// single-instance steps use the ordinary API and stay ordered with the
// ensemble's steps.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "chns.h"
#include "chns_ctx.hpp"
#include "chns_launch.hpp"
#include "chns_types.hpp"

struct chns_ensemble {
  struct Part {                 // a contiguous slice of a group, on its own stream
    int first = 0, count = 0;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
  };
  struct Group {
    int log2n = 0, model = 0, nph = 0, Jt = 0, split = 1;
    int sched = 0;              // FFT schedule of the batched passes (chns_batch.hip)
    std::vector<int> members;   // instance indices
    chns::Bufs* dB = nullptr;   // device arrays, one entry per member
    chns::KParams* dP = nullptr;
    std::vector<Part> parts;
  };
  int dev = 0;
  hipStream_t s0 = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<chns_ctx*> inst;
  std::vector<Group> groups;
  std::vector<void*> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graphs = true;           // false (CHNS_NO_GRAPH=1): every step launched directly
  std::string err;
};

namespace {

int efail(chns_ensemble* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}
int ehip(chns_ensemble* e, hipError_t r, const char* what) {
  return efail(e, CHNS_E_HIP, std::string(what) + ": " + hipGetErrorString(r));
}

#define ENS_HIP(e, call)                                  \
  do {                                                    \
    hipError_t r_ = (call);                               \
    if (r_ != hipSuccess) return ehip((e), r_, #call);    \
  } while (0)
#define ENS_LAUNCH(e, call)                                                 \
  do {                                                                      \
    int r_ = (call);                                                        \
    if (r_ != 0) return ehip((e), (hipError_t)r_, #call);                  \
  } while (0)

void drop_graph(chns_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

void drop_parts(chns_ensemble::Group& g) {
  for (auto& p : g.parts) {
    if (p.done) (void)hipEventDestroy(p.done);
    if (p.s) (void)hipStreamDestroy(p.s);
  }
  g.parts.clear();
}

// Split group g into `split` contiguous slices with a stream each.
int make_parts(chns_ensemble* e, chns_ensemble::Group& g, int split) {
  drop_parts(g);
  const int n = (int)g.members.size();
  split = std::max(1, std::min(split, n));
  g.split = split;
  for (int k = 0; k < split; ++k) {
    chns_ensemble::Part p;
    p.first = (int)((long)n * k / split);
    p.count = (int)((long)n * (k + 1) / split) - p.first;
    ENS_HIP(e, hipStreamCreate(&p.s));
    ENS_HIP(e, hipEventCreateWithFlags(&p.done, hipEventDisableTiming));
    g.parts.push_back(p);
  }
  return CHNS_OK;
}

// One ETD2RK step (both stages) of one slice, launched on its stream.
int step_part(chns_ensemble* e, const chns_ensemble::Group& g, const chns_ensemble::Part& p) {
  const chns::Bufs* B = g.dB + p.first;
  const chns::KParams* P = g.dP + p.first;
  const int L = g.log2n, M = g.model, nb = p.count;
  const int sc = g.sched;
  for (int st = 1; st <= 2; ++st) {
    if (g.nph > 0) ENS_LAUNCH(e, chns::launch_batch_rowB(L, M, B, P, nb, sc, p.s));
    ENS_LAUNCH(e, chns::launch_batch_colC(L, M, B, P, g.Jt, nb, sc, p.s));
    ENS_LAUNCH(e, chns::launch_batch_rowD(L, M, B, P, nb, sc, p.s));
    ENS_LAUNCH(e, chns::launch_batch_colE(L, M, st, B, P, 1 + g.nph, nb, sc, p.s));
  }
  return CHNS_OK;
}

// Schedules that differ from the default at grid size 2^log2n: the radix-16
// column passes from N = 256 up (8 points per thread below), half the rows
// per row workgroup below N = 2048 (one row per workgroup there).
bool schedule_distinct(int log2n, int sched) {
  if ((sched & 1) && log2n < 8) return false;
  if ((sched & 2) && log2n >= 11) return false;
  return sched >= 0 && sched < CHNS_ENSEMBLE_NSCHED;
}

// Steps per replay of the ensemble graph: as for chns_step, a graph of 8
// steps replaces 8 replays of a one-step graph.
constexpr int kGraphSteps = 8;

// One step of every group: s0 forks to the slice streams and joins back
// (launched directly, or recorded while s0 is captured).
int enqueue_step(chns_ensemble* e) {
  ENS_HIP(e, hipEventRecord(e->fork, e->s0));
  for (auto& g : e->groups)
    for (auto& p : g.parts) {
      ENS_HIP(e, hipStreamWaitEvent(p.s, e->fork, 0));
      const int rc = step_part(e, g, p);
      if (rc != CHNS_OK) return rc;
      ENS_HIP(e, hipEventRecord(p.done, p.s));
      ENS_HIP(e, hipStreamWaitEvent(e->s0, p.done, 0));
    }
  return CHNS_OK;
}

// Capture kGraphSteps steps of every group.
int build_graph(chns_ensemble* e) {
  drop_graph(e);
  ENS_HIP(e, hipStreamBeginCapture(e->s0, hipStreamCaptureModeThreadLocal));
  int rc = CHNS_OK;
  for (int k = 0; k < kGraphSteps && rc == CHNS_OK; ++k) rc = enqueue_step(e);
  hipGraph_t graph = nullptr;
  const hipError_t r2 = hipStreamEndCapture(e->s0, &graph);
  if (rc != CHNS_OK || r2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != CHNS_OK ? rc : ehip(e, r2, "hipStreamEndCapture");
  }
  e->graph = graph;
  ENS_HIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  ENS_HIP(e, hipGraphUpload(e->exec, e->s0));
  return CHNS_OK;
}

// Pass A (column iFFT of the phase fields) for instances whose row inputs are
// stale (after a state upload); the captured step assumes them current.
int refresh_inputs(chns_ensemble* e) {
  for (chns_ctx* c : e->inst) {
    if (!c->have_state) return efail(e, CHNS_E_STATE, "an instance has no state");
    if (c->nph > 0 && !c->hc_valid) {
      ENS_LAUNCH(e, chns::launch_colA(c->d.log2n, c->b, c->kp, c->nph, e->s0));
      c->hc_valid = true;
    }
  }
  return CHNS_OK;
}

// nsteps steps: kGraphSteps per graph replay, the remainder launched directly
// (every step, with graphs off).
int replay(chns_ensemble* e, int nsteps) {
  int n = 0;
  if (e->graphs) {
    if (!e->exec) {
      const int rc = build_graph(e);
      if (rc != CHNS_OK) return rc;
    }
    for (; n + kGraphSteps <= nsteps; n += kGraphSteps) ENS_HIP(e, hipGraphLaunch(e->exec, e->s0));
  }
  for (; n < nsteps; ++n) {
    const int rc = enqueue_step(e);
    if (rc != CHNS_OK) return rc;
  }
  return CHNS_OK;
}

}  // namespace

extern "C" const char* chns_ensemble_last_error(const chns_ensemble* e) {
  return e ? e->err.c_str() : "null ensemble";
}

extern "C" void chns_ensemble_destroy(chns_ensemble* e) {
  if (!e) return;
  (void)hipSetDevice(e->dev);
  if (e->s0) (void)hipStreamSynchronize(e->s0);
  drop_graph(e);
  for (auto& g : e->groups) drop_parts(g);
  for (chns_ctx* c : e->inst) {
    if (!c) continue;
    c->in_ensemble = false;
    chns_destroy(c);
  }
  for (void* q : e->allocs) (void)hipFree(q);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->s0) (void)hipStreamDestroy(e->s0);
  delete e;
}

extern "C" int chns_ensemble_create(chns_ensemble** out, const chns_params* p, int n, int device) {
  if (!out || !p || n < 1) return CHNS_E_ARG;
  *out = nullptr;
  hipError_t r = hipSetDevice(device);
  if (r != hipSuccess) {
    std::fprintf(stderr, "chns_ensemble_create: hipSetDevice(%d): %s\n", device, hipGetErrorString(r));
    return CHNS_E_HIP;
  }
  chns_ensemble* e = new chns_ensemble();
  e->dev = device;
  e->graphs = chns_graphs_default();
  auto bail = [&](int rc) {
    std::fprintf(stderr, "chns_ensemble_create: %s\n", e->err.c_str());
    chns_ensemble_destroy(e);
    return rc;
  };
  if ((r = hipStreamCreate(&e->s0)) != hipSuccess) return bail(ehip(e, r, "hipStreamCreate"));
  if ((r = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming)) != hipSuccess)
    return bail(ehip(e, r, "hipEventCreate"));
  std::map<std::pair<int, int>, int> gid;   // (log2n, model) -> group
  for (int b = 0; b < n; ++b) {
    if (p[b].gravity || p[b].vpm || p[b].sponge || p[b].smooth_filter || p[b].dealias != 0)
      return bail(efail(e, CHNS_E_ARG, "instance " + std::to_string(b) +
                                           ": the optional physics (gravity, vpm, sponge, smooth_filter) and the "
                                           "dealias rules other than 0 run on single contexts"));
    chns_ctx* c = nullptr;
    const int rc = chns_create_on(&c, &p[b], device, e->s0);
    if (rc != CHNS_OK) return bail(efail(e, rc, "instance " + std::to_string(b) + ": invalid parameters"));
    c->in_ensemble = true;
    e->inst.push_back(c);
    const auto key = std::make_pair(c->d.log2n, p[b].model);
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, (int)e->groups.size()).first;
      chns_ensemble::Group g;
      g.log2n = c->d.log2n;
      g.model = p[b].model;
      g.nph = c->nph;
      g.Jt = c->Jt;
      e->groups.push_back(g);
    }
    e->groups[it->second].members.push_back(b);
  }
  for (auto& g : e->groups) {
    const size_t m = g.members.size();
    std::vector<chns::Bufs> hb(m);
    std::vector<chns::KParams> hp(m);
    for (size_t k = 0; k < m; ++k) {
      hb[k] = e->inst[g.members[k]]->b;
      hp[k] = e->inst[g.members[k]]->kp;
    }
    void *qb = nullptr, *qp = nullptr;
    if ((r = hipMalloc(&qb, m * sizeof(chns::Bufs))) != hipSuccess) return bail(ehip(e, r, "hipMalloc"));
    e->allocs.push_back(qb);
    if ((r = hipMalloc(&qp, m * sizeof(chns::KParams))) != hipSuccess) return bail(ehip(e, r, "hipMalloc"));
    e->allocs.push_back(qp);
    g.dB = (chns::Bufs*)qb;
    g.dP = (chns::KParams*)qp;
    if ((r = hipMemcpy(qb, hb.data(), m * sizeof(chns::Bufs), hipMemcpyHostToDevice)) != hipSuccess ||
        (r = hipMemcpy(qp, hp.data(), m * sizeof(chns::KParams), hipMemcpyHostToDevice)) != hipSuccess)
      return bail(ehip(e, r, "hipMemcpy"));
    const int rc = make_parts(e, g, 1);
    if (rc != CHNS_OK) return bail(rc);
  }
  *out = e;
  return CHNS_OK;
}

extern "C" int chns_ensemble_size(const chns_ensemble* e) { return e ? (int)e->inst.size() : 0; }

extern "C" chns_ctx* chns_ensemble_instance(chns_ensemble* e, int b) {
  if (!e || b < 0 || b >= (int)e->inst.size()) return nullptr;
  return e->inst[b];
}

extern "C" int chns_ensemble_step(chns_ensemble* e, int nsteps) {
  if (!e) return CHNS_E_ARG;
  if (nsteps < 0) return efail(e, CHNS_E_ARG, "nsteps < 0");
  if (nsteps == 0) return CHNS_OK;
  ENS_HIP(e, hipSetDevice(e->dev));
  const int rc = refresh_inputs(e);
  if (rc != CHNS_OK) return rc;
  return replay(e, nsteps);
}

extern "C" int chns_ensemble_sync(chns_ensemble* e) {
  if (!e) return CHNS_E_ARG;
  ENS_HIP(e, hipSetDevice(e->dev));
  ENS_HIP(e, hipStreamSynchronize(e->s0));
  return CHNS_OK;
}

extern "C" int chns_ensemble_tune(chns_ensemble* e, int reps, int* splits) {
  // Greedy per group, largest grids first: time `reps` replays of the graph
  // (kGraphSteps steps each) for every FFT schedule of the group's batched
  // passes that differs from the default at its size, each with the group
  // split into 1, 2 and 4 concurrently launched slices, and keep the fastest
  // pair (the other groups keep their current choice meanwhile).  The
  // instances' spectral state is saved before and restored after, so tuning
  // leaves the physics untouched.
  if (!e || reps < 1) return CHNS_E_ARG;
  ENS_HIP(e, hipSetDevice(e->dev));
  // queued replays finish before their graph and slice streams are replaced
  ENS_HIP(e, hipStreamSynchronize(e->s0));
  int rc = refresh_inputs(e);
  if (rc != CHNS_OK) return rc;
  std::vector<void*> save;
  auto free_save = [&] {
    for (void* q : save) (void)hipFree(q);
  };
  for (chns_ctx* c : e->inst) {
    const size_t bytes = (size_t)c->d.N * c->d.H * sizeof(chns::cd);
    chns::cd* f[3] = {c->b.W, c->b.C1, c->b.C2};
    for (int k = 0; k <= c->nph; ++k) {
      void* q = nullptr;
      hipError_t r = hipMalloc(&q, bytes);
      if (r != hipSuccess) {
        free_save();
        return ehip(e, r, "hipMalloc(tune)");
      }
      save.push_back(q);
      r = hipMemcpyAsync(q, f[k], bytes, hipMemcpyDeviceToDevice, e->s0);
      if (r != hipSuccess) {
        free_save();
        return ehip(e, r, "hipMemcpyAsync(tune)");
      }
    }
  }
  std::vector<int> order(e->groups.size());
  for (size_t g = 0; g < order.size(); ++g) order[g] = (int)g;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const auto& A = e->groups[a];
    const auto& B = e->groups[b];
    return (double)A.members.size() * (1 << (2 * A.log2n)) > (double)B.members.size() * (1 << (2 * B.log2n));
  });
  auto time_graph = [&](double* ms) -> int {
    drop_graph(e);   // the slices changed: replay() captures them anew
    int r;
    if ((r = replay(e, kGraphSteps)) != CHNS_OK) return r;   // warm-up: one graph replay
    ENS_HIP(e, hipStreamSynchronize(e->s0));
    const auto t0 = std::chrono::steady_clock::now();
    if ((r = replay(e, reps * kGraphSteps)) != CHNS_OK) return r;   // reps graph replays
    ENS_HIP(e, hipStreamSynchronize(e->s0));
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
          (reps * kGraphSteps);
    return CHNS_OK;
  };
  for (int gi : order) {
    auto& g = e->groups[gi];
    int best = 1, best_sched = 0;
    double best_ms = 1e300;
    for (int sched = 0; sched < CHNS_ENSEMBLE_NSCHED; ++sched) {
      if (sched && !schedule_distinct(g.log2n, sched)) continue;
      g.sched = sched;
      for (int split : {1, 2, 4}) {
        if (split > (int)g.members.size()) break;
        double ms = 0.0;
        if ((rc = make_parts(e, g, split)) != CHNS_OK || (rc = time_graph(&ms)) != CHNS_OK) {
          free_save();
          return rc;
        }
        if (ms < best_ms) {
          best_ms = ms;
          best = split;
          best_sched = sched;
        }
      }
    }
    g.sched = best_sched;
    if ((rc = make_parts(e, g, best)) != CHNS_OK) {
      free_save();
      return rc;
    }
  }
  drop_graph(e);
  size_t k = 0;
  for (chns_ctx* c : e->inst) {
    const size_t bytes = (size_t)c->d.N * c->d.H * sizeof(chns::cd);
    chns::cd* f[3] = {c->b.W, c->b.C1, c->b.C2};
    for (int q = 0; q <= c->nph; ++q) {
      hipError_t r = hipMemcpyAsync(f[q], save[k++], bytes, hipMemcpyDeviceToDevice, e->s0);
      if (r != hipSuccess) {
        free_save();
        return ehip(e, r, "hipMemcpyAsync(restore)");
      }
    }
    c->hc_valid = false;
  }
  hipError_t r = hipStreamSynchronize(e->s0);
  free_save();
  if (r != hipSuccess) return ehip(e, r, "hipStreamSynchronize");
  if (splits)
    for (size_t g = 0; g < e->groups.size(); ++g) splits[g] = e->groups[g].split;
  return CHNS_OK;
}

extern "C" int chns_ensemble_groups(const chns_ensemble* e, int* log2n, int* model, int* count) {
  if (!e) return 0;
  for (size_t g = 0; g < e->groups.size(); ++g) {
    if (log2n) log2n[g] = e->groups[g].log2n;
    if (model) model[g] = e->groups[g].model;
    if (count) count[g] = (int)e->groups[g].members.size();
  }
  return (int)e->groups.size();
}

extern "C" int chns_ensemble_schedule(const chns_ensemble* e, int* split, int* sched) {
  if (!e) return 0;
  for (size_t g = 0; g < e->groups.size(); ++g) {
    if (split) split[g] = e->groups[g].split;
    if (sched) sched[g] = e->groups[g].sched;
  }
  return (int)e->groups.size();
}

extern "C" int chns_ensemble_set_schedule(chns_ensemble* e, int group, int split, int sched) {
  if (!e) return CHNS_E_ARG;
  if (group < 0 || group >= (int)e->groups.size()) return efail(e, CHNS_E_ARG, "no such group");
  if (split < 1 || sched < 0 || sched >= CHNS_ENSEMBLE_NSCHED)
    return efail(e, CHNS_E_ARG, "split must be >= 1 and sched in [0, CHNS_ENSEMBLE_NSCHED)");
  ENS_HIP(e, hipSetDevice(e->dev));
  ENS_HIP(e, hipStreamSynchronize(e->s0));   // the slices' streams are replaced
  auto& g = e->groups[group];
  g.sched = sched;
  drop_graph(e);
  return make_parts(e, g, split);
}

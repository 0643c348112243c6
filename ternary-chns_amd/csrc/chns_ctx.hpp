// Context of the C ABI (internal to the library).
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "chns.h"
#include "chns_host.hpp"
#include "chns_types.hpp"

struct chns_ctx {
  chns_params p{};
  chns_host::Derived d{};
  int dev = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;                    // false: an ensemble's stream
  bool in_ensemble = false;                  // owned by a chns_ensemble
  int Jt = 0, ldJ = 0, ldF = 0, nph = 0;
  chns::KParams kp{};
  chns::Bufs b{};
  std::vector<void*> allocs;
  std::vector<int> amask;
  bool have_state = false, hc_valid = false;
  std::string err;
  // captured steps of chns_step, replayed together (buffers and parameters
  // are fixed when the context is created, so the graph stays valid)
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  bool graphs = true;                        // chns_graphs_enable / CHNS_NO_GRAPH
  // divergence status (chns_status): the device flag b.flag raised by pass E,
  // and the host-side free-energy guard of the last chns_free_energy
  int energy_flag = 0;
  // event timing (chns_timing_enable)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_kind;   // one entry per recorded pair
  size_t ev_used = 0;
  // diagnostics scratch (allocated on first use)
  chns::cd *dT1 = nullptr, *dT2 = nullptr;   // tiled [N][ldF] column-iFFT outputs
  double* dred = nullptr;                    // reduction partials + result
  double* dspec = nullptr;                   // spectrum row partials
  double* dspec_out = nullptr;               // spectrum sums
  chns::cd* dstage = nullptr;                // one field in the IDXC layout (state transfer)
  struct chns_adj_state* adj = nullptr;      // discrete-adjoint storage (chns_adjoint.cpp)
};

int chns_fail(chns_ctx* c, int code, const std::string& msg);
int chns_hip_fail(chns_ctx* c, hipError_t e, const char* what);
int chns_dalloc_bytes(chns_ctx* c, void** p, size_t bytes);
// chns_create on a caller-owned stream (nullptr: the context creates its own)
int chns_create_on(chns_ctx** out, const chns_params* p, int device, hipStream_t shared);
// c->dstage: one field in the reference layout IDXC [N][H], which
// chns::launch_transpose turns into or out of the column-major [H][N] state
int chns_stage_buffer(chns_ctx* c);
// CHNS_NO_GRAPH=1 in the environment: graphs off by default
bool chns_graphs_default();
// packed real [N][N] c2r (scaled 1/(Nx*Ny)) of a column-major spectral field S
// of this context, copied to host memory out (synchronises)
int chns_real_of(chns_ctx* c, const chns::cd* S, double* out);
// diagnostics scratch (dT1, dT2, dred), allocated on first use
int chns_diag_prepare(chns_ctx* c);
// the real fields of the free energy in the row-pass buffers: c1, c2, d/dx c1,
// d/dy c1, d/dx c2, d/dy c2 (binary: c, d/dx c, d/dy c, then nulls), of the
// state or of its truncation
int chns_fe_fields(chns_ctx* c, bool trunc, double** f);
// the two sums of free_energy.c (bulk, gradient) over those fields with the
// weights g1..g3 (ternary), unscaled, into q[0..1] (synchronises)
int chns_fe_sums(chns_ctx* c, const double* const* f, double g1, double g2, double g3, double* q);
// frees the adjoint storage (chns_destroy, chns_adj_release)
void chns_adj_free(chns_ctx* c);
// device address of chns_free_energy_noalloc's static scratch array on `dev`
// (current device), looked up once per device (chns_noalloc.hip)
int chns_noalloc_setup(int dev, double** out);

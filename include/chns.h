/*
 * chns.h - C ABI of the MI355X-native CHNS-2D ETD2RK hot path.
 *
 * Drop-in boundary for the time loop of the reference pseudo-spectral
 * Cahn-Hilliard-Navier-Stokes solver (phynadia/Ternary-CHNS, "CHNS-2D-MPI").
 * Plain C types only: pointers, sizes, doubles, ints.  Complex arrays are
 * interleaved (re, im) doubles with the reference layout IDXC(i,j) = j + Ny2*i
 * (global.h:13), i.e. [Nx][Ny/2+1] complex, x slow.  Real arrays are packed
 * [Nx][Ny] doubles (the reference's IDXR rows are [Ny+2] with two pad
 * doubles, global.h:14; callers strip the pad).
 *
 * Every call returns 0 on success and a negative CHNS_E* code on failure;
 * nothing aborts.  chns_last_error() gives the message.  A context owns one
 * HIP device and one stream; calls on a context are stream-ordered and a
 * context is not thread-safe.  State lives on the device between calls.
 */
#ifndef CHNS_H
#define CHNS_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHNS_OK 0
#define CHNS_E_ARG (-1)     /* invalid argument or unsupported configuration */
#define CHNS_E_HIP (-2)     /* HIP runtime error */
#define CHNS_E_STATE (-3)   /* call out of order (e.g. no state set) */
#define CHNS_E_DIVERGED (-4) /* chns_status: the state went non-finite or the free-energy guard tripped */

/* models: macro.h:7-8 (neither BINARY nor TERNARY = Navier-Stokes only) */
#define CHNS_MODEL_NS 0
#define CHNS_MODEL_BINARY 1
#define CHNS_MODEL_TERNARY 2

/* Runtime form of the reference's compile-time parameters (global.h:20-33,
 * 102-104, 127-134).  `lambda` is the mobility.  epsilon <= 0 selects the
 * reference default 3*dx.  f_amp/f_k: Kolmogorov forcing f_amp*sin(f_k*y) on
 * the vorticity (new: the reference forcing.c does not compile).
 * Optional physics (macro.h switches): gravity - buoyancy torque
 * i kx F(-g (rho(c) - 1)) of nonlin.c:107-171 with rho1..rho3, g of
 * global.h:106-110,135-140 (BINARY uses rho1, rho2); vpm - penalised walls
 * chi(x) of init.c:184-206 with eta = 1.5 dt (global.h:47-48) acting as in
 * nonlin.c:63-95,209-211,294-296,337-339 and surface_tension.c:83-85,
 * 151-154,202-205; sponge - the sponge(y)/eta damping of the vorticity
 * (init.c:207-211, nonlin.c:50-99), which needs vpm.
 * dealias: the truncation() rule.  0 (default) is the reference's circular
 * cut, zero where kx^2 + ky^2 > ((Nx/4) 2pi/lx)^2 (functions.c:56-59,75-78);
 * 1 is the 2/3 rule, zero where |kx| > (Nx/3) 2pi/lx or |ky| > (Ny/3) 2pi/ly;
 * 2 is the circular cut at Nx/3 instead of Nx/4 (the Nx/3 radius of
 * energy.c:13-15).  Rules 1 and 2 run the unpruned kernels of the smooth
 * filter with a 0/1 factor table; single contexts only, not with
 * smooth_filter (which replaces the cut). */
typedef struct {
  int Nx, Ny, model;
  double lx, ly, nu, dt;
  double sigma, epsilon, lambda;
  double sigma_12, sigma_13, sigma_23, Xlambda;
  int smooth_filter; /* macro.h:15 SMOOTH_FILTER: truncation() multiplies by exp(-36 pp^18) twice
                        instead of cutting (functions.c:70-73); single contexts only */
  double f_amp;
  int f_k;
  int gravity;       /* macro.h GRAVITY */
  double rho1, rho2, rho3, g;
  int vpm;           /* macro.h VPM */
  int sponge;        /* macro.h SPONGE (requires vpm) */
  int dealias;       /* truncation rule: 0 reference N/4 circle, 1 2/3 rule (square), 2 N/3 circle */
} chns_params;

typedef struct chns_ctx chns_ctx;

/* Fill p with the reference defaults (global.h, macro.h) for `model`. */
void chns_params_default(chns_params *p, int model);

/* Allocate a context on HIP device `device`.  Replaces main.c:37-194
 * (MPI/FFTW init, allocation, plans) and kxky() (functions.c:11-32). */
int chns_create(chns_ctx **out, const chns_params *p, int device);
void chns_destroy(chns_ctx *ctx);
const char *chns_last_error(const chns_ctx *ctx);

/* Upload / download the spectral state (host arrays, IDXC layout, 2*Nx*Ny2
 * doubles each).  cphi1/cphi2 are ignored (may be NULL) when the model has no
 * such field; for BINARY cphi1 is the reference's cphi.  Once a state is
 * set, a NULL array in chns_set_state leaves that field as it is on the
 * device.  Replaces the state arrays cwz, cphi1, cphi2 of global.c:13-49. */
int chns_set_state(chns_ctx *ctx, const double *cwz, const double *cphi1, const double *cphi2);
int chns_get_state(chns_ctx *ctx, double *cwz, double *cphi1, double *cphi2);

/* Initial state from packed real [Nx][Ny] fields: dfft + truncation, as
 * init.c:253-262 (INIT_READ_DATA) and init.c:291-294.  wz may be NULL (zero
 * vorticity, init.c:78-100). */
int chns_set_real_state(chns_ctx *ctx, const double *wz, const double *phi1, const double *phi2);

/* Advance nsteps ETD2RK steps on the device: etd2rk() (etd2rk.c:14-224)
 * called nsteps times, including potential() (chem_potential.c:12-94),
 * nonlin() (nonlin.c:13-381), surface_tension() (surface_tension.c:13-227),
 * get_velocity() (velocity.c:3-44), curl() (curl.c:16-31), truncation()
 * (functions.c:47-83) and the dfft/ifft pairs (fft.c:11-31) inside them.
 * Returns when the work is queued; chns_sync() waits.  Steps run 8 at a
 * time as replays of a HIP graph of 8 captured steps (captured and uploaded
 * on the first call, whatever its nsteps), the remainder as direct launches;
 * with event timing on (chns_timing_enable), with graphs off
 * (chns_graphs_enable), or while the caller captures the context's stream
 * into a graph of its own, every kernel is launched directly.  Both paths
 * compute bit-identical states. */
int chns_step(chns_ctx *ctx, int nsteps);
int chns_sync(chns_ctx *ctx);

/* HIP graph replays of chns_step on (1, default) or off (0: every kernel
 * launched directly, e.g. for profilers that trace launches one by one).
 * The environment variable CHNS_NO_GRAPH=1 makes 0 the default of every
 * context and ensemble created while it is set.  No reference counterpart. */
int chns_graphs_enable(chns_ctx *ctx, int enable);

/* Divergence status.  The reference's only guard is free_energy.c:151-154,
 * MPI_Abort when the bulk free energy ephi >= 100; this library never
 * aborts and reports instead.  Waits for the context's queued work, then
 * returns CHNS_E_DIVERGED if a step since the state was last set wrote a
 * non-finite value into it (a flag raised on the device by the ETD2RK
 * update, CHNS_FLAG_NONFINITE), or if the last chns_free_energy found
 * ephi >= 100 or non-finite (CHNS_FLAG_ENERGY); CHNS_OK otherwise.  flags
 * (may be NULL) receives the bits.  chns_set_state and chns_set_real_state
 * clear both. */
#define CHNS_FLAG_NONFINITE 1
#define CHNS_FLAG_ENERGY 2
int chns_status(chns_ctx *ctx, int *flags);

/* In-place truncation of the phase fields, the state side effect of
 * spectrum() (spectrum.c:22-27) called at main.c:226,281-284.  The next
 * chns_step re-reads the truncated fields (with SMOOTH_FILTER truncation is
 * not idempotent, so the stage inputs are rebuilt from the new state). */
int chns_truncate_state(chns_ctx *ctx);

/* ---- diagnostics (device computation, fixed-order reductions) ---------- */

/* Bulk and gradient free energy of the untruncated phase fields
 * (free_energy.c:13-157); BINARY and TERNARY only. */
int chns_free_energy(chns_ctx *ctx, double *ephi, double *egphi);
/* chns_free_energy without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * or time loops, harnesses that forbid hipMalloc).  Same outputs, bit for
 * bit, and the same free-energy guard for chns_status.  Its reduction
 * partials live in a static device array of 41 KB per GPU (N <= 2048; larger
 * N returns CHNS_E_ARG) and its transforms reuse the context's step buffers,
 * which hold nothing between steps.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (two threads or streams) race. */
int chns_free_energy_noalloc(chns_ctx *ctx, double *ephi, double *egphi);
/* Mean kinetic energy eu and the per-phase kinetic energies averaged over
 * the points with c_i >= 0.5 (kinetic_energy.c:3-144); eph2 is ternary only,
 * eph1/eph2 untouched for NS.  Any output may be NULL. */
int chns_kinetic_energy(chns_ctx *ctx, double *eu, double *eph1, double *eph2);
/* chns_kinetic_energy with the per-phase point counts n1, n2 (the points
 * with c_i >= 0.5: count, count1, count2 of kinetic_energy.c:55-80), which
 * decide whether the reference writes its BINARY line (kinetic_energy.c:
 * 103-108 prints only when count != 0).  Any output may be NULL. */
int chns_kinetic_energy_counts(chns_ctx *ctx, double *eu, double *eph1, double *eph2, double *n1, double *n2);
/* Interface counts of the phase fields thresholded at 0.5 (perimeter.c:127-238):
 * counts[0] sign changes to the x+1/y+1 neighbour and counts[1] points inside
 * for phase 1, counts[2..3] for phase 2 (ternary).  The reference prints
 * Sp_t = count*dx, Sp_0 = 2 sqrt(pi inside dx dy), def = Sp_t/Sp_0 - 1. */
int chns_perimeter(chns_ctx *ctx, double *counts);
/* Shell spectra of spectrum.c:16-121: ek, ensk and the phase spectra sph1
 * (binary: sph), sph2, sp12 have nshell entries; ekx, eky have Nx/2+2 (the
 * reference's kxm-sized arrays are too short for factor <= 1).  NULL skips an
 * output.  truncate_state != 0 also applies spectrum()'s truncation of the
 * phase fields to the state (spectrum.c:22-27). */
int chns_spectrum(chns_ctx *ctx, int truncate_state, double *ek, double *ensk, double *sph1, double *sph2,
                  double *sp12, double *ekx, double *eky);
/* Packed real [Nx][Ny] field which = 0 (vorticity), 1, 2 (phase fields),
 * truncated first when truncate != 0 (the DUMP_REAL_MPI path, main.c:332-358). */
int chns_real_field(chns_ctx *ctx, int which, int truncate, double *out);
/* Lengths of the spectrum outputs (nshell = round(Nx/sqrt(2) + 1), main.c:60). */
int chns_diag_sizes(const chns_params *p, int *nshell, int *nxy);

/* Per-pass kernel timing with HIP events recorded on the context's stream
 * around every launch (measurement aid, off by default; no reference
 * counterpart - the reference only times file output with MPI_Wtime,
 * main.c:298-299,362-363).  Pass kinds: */
#define CHNS_PASS_COLA 0   /* dealiased column iFFT of the phase fields */
#define CHNS_PASS_ROWB 1   /* row c2r -> chemical potential -> row r2c */
#define CHNS_PASS_COLC 2   /* column FFT/iFFTs of mu, velocity, gradients */
#define CHNS_PASS_ROWD 3   /* row c2r -> nonlinear products -> row r2c */
#define CHNS_PASS_COLE1 4  /* column FFT + ETD2RK stage 1 */
#define CHNS_PASS_COLE2 5  /* column FFT + ETD2RK stage 2 */
#define CHNS_NPASS 6
int chns_timing_enable(chns_ctx *ctx, int enable);
/* Synchronises, then returns the summed kernel milliseconds and launch
 * counts per pass kind since the last read, and resets them. */
int chns_timing_read(chns_ctx *ctx, double *total_ms, int *launches);

/* ---- ensembles (no reference counterpart: the reference runs one grid per
 * MPI job; these run many independent instances on one GPU) -------------- */

typedef struct chns_ensemble chns_ensemble;

/* n independent instances, instance b with parameters p[b] (any mix of grid
 * sizes and models).  Instances of equal size and model form a group whose
 * passes run as one batched launch each; groups run concurrently. */
int chns_ensemble_create(chns_ensemble **out, const chns_params *p, int n, int device);
void chns_ensemble_destroy(chns_ensemble *e);
const char *chns_ensemble_last_error(const chns_ensemble *e);
int chns_ensemble_size(const chns_ensemble *e);
/* The context of instance b (NULL if out of range): state transfer,
 * diagnostics and single-instance steps through the ordinary calls above.
 * Its work runs on the ensemble's stream, ordered with chns_ensemble_step.
 * Owned by the ensemble: chns_destroy on it does nothing. */
chns_ctx *chns_ensemble_instance(chns_ensemble *e, int b);
/* nsteps ETD2RK steps of every instance (each instance needs a state).  Eight
 * steps of the whole ensemble are captured into a HIP graph on first use and
 * replayed per eight steps, the remainder launched directly.  Returns when
 * queued; chns_ensemble_sync waits. */
int chns_ensemble_step(chns_ensemble *e, int nsteps);
int chns_ensemble_sync(chns_ensemble *e);
/* FFT schedules of a group's batched passes: bit 0 runs the column passes
 * with 16 points per thread (radix-16 stages, one LDS exchange fewer per
 * 2048-point transform) instead of 8, bit 1 gives the row passes half the
 * rows per workgroup.  0 is the single-context schedule.  All schedules give
 * the same results to round-off. */
#define CHNS_ENSEMBLE_NSCHED 4
/* Autotune every group (largest grids first): for each FFT schedule that
 * differs from the default at the group's grid size and each launch split
 * (1, 2 or 4 concurrent slices), time `reps` replays of the 8-step graph,
 * and keep the fastest; the instances' states are restored afterwards.
 * splits (may be NULL) receives the chosen split per group, in group order;
 * chns_ensemble_schedule gives both choices. */
int chns_ensemble_tune(chns_ensemble *e, int reps, int *splits);
/* Current split and FFT schedule per group (arrays may be NULL); returns the
 * number of groups. */
int chns_ensemble_schedule(const chns_ensemble *e, int *split, int *sched);
/* Set a group's split and schedule by hand (waits for queued work). */
int chns_ensemble_set_schedule(chns_ensemble *e, int group, int split, int sched);
/* Groups in order: log2 N, model and instance count (arrays may be NULL);
 * returns the number of groups. */
int chns_ensemble_groups(const chns_ensemble *e, int *log2n, int *model, int *count);

/* ---- discrete adjoint of the ternary step (no reference counterpart: the
 * reference only integrates forward; this differentiates etd2rk() of
 * etd2rk.c:14-224 with everything it calls) ------------------------------
 *
 * Usage: set the initial state, chns_adj_forward (the state ends at the final
 * time), seed dL/d(final state) with chns_adj_seed_spectrum,
 * chns_adj_seed_free_energy or chns_adj_seed,
 * then chns_adj_backward.  Gradients are with respect to the real fields that
 * chns_set_real_state turns into the initial state, and to sigma_12/13/23.
 * Spectral adjoints follow one convention: Xbar is the unnormalised r2c of
 * the real-space gradient dL/dx of x = ifft(X) (fft.c:16-31 scaling), so
 * dL = sum_x xbar dx.  TERNARY only; storage lives on the device until
 * chns_adj_release or chns_destroy. */

/* Advance nsteps ETD2RK steps from the current state like chns_step, storing
 * the state every ckpt_every steps on the device (checkpoints for the reverse
 * sweep, which replays one segment of ckpt_every steps at a time). */
int chns_adj_forward(chns_ctx *ctx, int nsteps, int ckpt_every);
/* Loss = sum_s (sph1(s) - target(s))^2 over the nshell shells of the phase-1
 * spectrum of spectrum.c:83-103 at the current (final) state, and its seed.
 * loss and sph1 (nshell entries) may be NULL.  The state is not truncated. */
int chns_adj_seed_spectrum(chns_ctx *ctx, const double *target, double *loss, double *sph1);
/* Loss = total free energy ephi + egphi (free_energy.c:13-157) of the
 * truncated final phase fields T(c1), T(c2) (truncation() as every read of
 * the state inside the step applies it; the reference's diagnostic reads the
 * untruncated fields, whose extra modes the step never reads back), and its
 * seed; its direct dependence on sigma_ij through the gamma_i weights enters
 * g_sigma.  loss, ephi, egphi may be NULL.  The state is not changed. */
int chns_adj_seed_free_energy(chns_ctx *ctx, double *loss, double *ephi, double *egphi);
/* Seed with caller-given spectral adjoints of the final state (IDXC layout,
 * NULL = zero); they are truncated to the kept modes. */
int chns_adj_seed(chns_ctx *ctx, const double *wbar, const double *c1bar, const double *c2bar);
/* Reverse sweep.  g_wz, g_phi1, g_phi2: packed real [Nx][Ny] gradients with
 * respect to the initial real fields; g_sigma[3]: dL/d(sigma_12, sigma_13,
 * sigma_23).  Any output may be NULL.  The context state is unchanged. */
int chns_adj_backward(chns_ctx *ctx, double *g_wz, double *g_phi1, double *g_phi2, double *g_sigma);
/* After chns_adj_backward: spectral adjoints of the initial state (IDXC) and
 * the 15 coefficient partial sums behind g_sigma (fac, G1..3, 1/g1..3, mc1,
 * mc2, A11, A12, A21, A22, lambda/g1, lambda/g2).  Any output may be NULL. */
int chns_adj_get(chns_ctx *ctx, double *wbar, double *c1bar, double *c2bar, double *partials);
/* Free the adjoint's device storage (checkpoints, trajectory, work arrays). */
int chns_adj_release(chns_ctx *ctx);
/* store != 0 (default): the replay of a segment keeps the forward
 * intermediates of its 2*ckpt_every stages (at 2048^2 about 236 MB each) when
 * they fit in the free device memory, so the reverse sweep does not recompute
 * them; 0: always recompute.  Takes effect at the next chns_adj_forward; the
 * gradients agree to round-off (bitwise below 256^2). */
int chns_adj_set_store(chns_ctx *ctx, int store);

/* ---- host-only helpers (no device needed) ---------------------------- */

/* ETD2RK coefficients per mode, IDXC order [Nx][Ny2] each, in the exact
 * operation order of etd2rk.c:88-127 and :162-203 with the C library exp.
 * beta/z1c/z2c are left untouched for CHNS_MODEL_NS (may be NULL). */
int chns_etd_tables(const chns_params *p, double *eta, double *z1w, double *z2w,
                    double *beta, double *z1c, double *z2c);

/* Dealias extent per ky column j < Ny/2+1: the largest |kx index| kept by
 * truncation() (functions.c:47-83) under p->dealias, -1 if the column is
 * fully truncated. */
int chns_dealias_extent(const chns_params *p, int *amax);

/* SMOOTH_FILTER factor exp(-36 pp^18), pp = ksqr/ksqrmax, per mode in IDXC
 * order [Nx][Ny2] (functions.c:54,70-73, C library exp and pow); truncation()
 * multiplies by it twice.  Defined whatever p->smooth_filter is. */
int chns_filter_table(const chns_params *p, double *ft);

#ifdef __cplusplus
}
#endif
#endif /* CHNS_H */

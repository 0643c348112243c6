/*
 * chns2d - time-loop driver of the MI355X CHNS-2D step: main.c written over
 * the reference-named entry points of chns_ref.h (etd2rk, free_energy,
 * kinetic_energy, perimeter, spectrum), which call the C ABI of chns.h.
 *
 * Mirrors the reference main.c:24-368 on one GPU: initial condition
 * (init.c), t = 0 diagnostics and dump, then for it = 1..maxiter:
 * t += dt, one ETD2RK step, free_energy / kinetic_energy / perimeter every
 * `diag_every` steps (1 = every step, as main.c:276-279), spectrum every
 * navg/2 steps with its truncation of the phase fields (main.c:281-284), and
 * a truncated real dump every navg steps (DUMP_REAL_MPI, main.c:332-358).
 * Output files and printf formats are the reference's:
 *   data/ken_t.dat       kinetic_energy.c:109,135,141
 *   data/free_en_t.dat   free_energy.c:149 (exit 3 when chns_status reports ephi >= 100
 *                        or a non-finite state, as free_energy.c:151-154 aborts)
 *   data/gamma.dat | data/gamma1.dat, data/gamma2.dat   perimeter.c:123,264-265
 *   data/volume_*.dat    opened and left empty, as main.c:123,161
 *   spectra/{spectrum,Xspectrum,Yspectrum}.N           spectrum.c:150-187
 *   vort/wz.N, phase_field/phi[12].N (raw float64 [Nx][Ny])  write_array.c:19-79
 *
 * Parameters are key=value arguments (names of global.h / macro.h):
 *   model=ns|binary|ternary Nx lx nu dt maxiter nfile sigma epsilon lambda
 *   sigma_12 sigma_13 sigma_23 Xlambda f_amp f_k
 *   gravity rho1 rho2 rho3 g vpm sponge   (macro.h GRAVITY / VPM / SPONGE switches)
 *   smooth_filter                         (macro.h SMOOTH_FILTER)
 *   dealias=0|1|2  truncation rule: reference N/4 circle, 2/3 rule, N/3 circle
 *   init=noise|read|droplet|zero  nranks (noise: srand(10+2*rank) per slab)
 *   read1=path read2=path (raw float64 [Nx][Ny], INIT_READ_DATA; defaults
 *   init_lens/phi1.in, init_lens/phi2.in, binary init_field/phi.in)
 *   diag_every=K  out=DIR  device=D
 *   mirror=1: the host arrays hold the state across every call (the
 *   reference's in-place semantics; see chns_ref.h), else device-resident
 *   ic_only=1: write the initial real fields to init_field/ic{1,2}.bin and
 *   stop before any device call
 * Defaults are the reference defaults (TERNARY, N = 1024, L = 4 pi, ...).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "chns.h"
#include "chns_ref.h"

static int read_raw(const char *path, double *u, size_t n) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  size_t got = fread(u, sizeof(double), n, f);   /* read_array.c:21-31: Nx*Ny doubles, row-major */
  fclose(f);
  return got == n ? 0 : -1;
}

static void write_raw(const char *path, const double *u, size_t n) {
  FILE *f = fopen(path, "wb");
  if (!f) {
    fprintf(stderr, "chns2d: cannot write %s\n", path);
    exit(1);
  }
  fwrite(u, sizeof(double), n, f);
  fclose(f);
}

int main(int argc, char **argv) {
  chns_params p;
  int model = CHNS_MODEL_TERNARY;
  for (int a = 1; a < argc; ++a)
    if (!strncmp(argv[a], "model=", 6))
      model = !strcmp(argv[a] + 6, "ns") ? CHNS_MODEL_NS : (!strcmp(argv[a] + 6, "binary") ? CHNS_MODEL_BINARY : CHNS_MODEL_TERNARY);
  chns_params_default(&p, model);
  double maxiter = 2e5, nfile = 1000;          /* global.h:32-33 */
  const char *init = model == CHNS_MODEL_BINARY ? "droplet" : (model == CHNS_MODEL_TERNARY ? "noise" : "zero");
  const char *read1 = model == CHNS_MODEL_BINARY ? "init_field/phi.in" : "init_lens/phi1.in", *read2 = "init_lens/phi2.in", *out = ".";
  int nranks = 1, diag_every = 1, device = 0, ic_only = 0, mirror = 0;
  for (int a = 1; a < argc; ++a) {
    char *eq = strchr(argv[a], '=');
    if (!eq) continue;
    *eq = 0;
    const char *k = argv[a], *v = eq + 1;
    if (!strcmp(k, "Nx") || !strcmp(k, "N")) p.Nx = p.Ny = atoi(v);
    else if (!strcmp(k, "lx") || !strcmp(k, "L")) p.lx = p.ly = atof(v);
    else if (!strcmp(k, "nu")) p.nu = atof(v);
    else if (!strcmp(k, "dt")) p.dt = atof(v);
    else if (!strcmp(k, "maxiter")) maxiter = atof(v);
    else if (!strcmp(k, "nfile")) nfile = atof(v);
    else if (!strcmp(k, "sigma")) p.sigma = atof(v);
    else if (!strcmp(k, "epsilon")) p.epsilon = atof(v);
    else if (!strcmp(k, "lambda")) p.lambda = atof(v);
    else if (!strcmp(k, "sigma_12")) p.sigma_12 = atof(v);
    else if (!strcmp(k, "sigma_13")) p.sigma_13 = atof(v);
    else if (!strcmp(k, "sigma_23")) p.sigma_23 = atof(v);
    else if (!strcmp(k, "Xlambda")) p.Xlambda = atof(v);
    else if (!strcmp(k, "f_amp")) p.f_amp = atof(v);
    else if (!strcmp(k, "f_k")) p.f_k = atoi(v);
    else if (!strcmp(k, "gravity")) p.gravity = atoi(v);   /* macro.h GRAVITY */
    else if (!strcmp(k, "rho1")) p.rho1 = atof(v);
    else if (!strcmp(k, "rho2")) p.rho2 = atof(v);
    else if (!strcmp(k, "rho3")) p.rho3 = atof(v);
    else if (!strcmp(k, "g")) p.g = atof(v);
    else if (!strcmp(k, "vpm")) p.vpm = atoi(v);           /* macro.h VPM */
    else if (!strcmp(k, "sponge")) p.sponge = atoi(v);     /* macro.h SPONGE */
    else if (!strcmp(k, "smooth_filter")) p.smooth_filter = atoi(v);   /* macro.h SMOOTH_FILTER */
    else if (!strcmp(k, "dealias")) p.dealias = atoi(v);   /* truncation rule (0 = functions.c:56-59) */
    else if (!strcmp(k, "init")) init = v;
    else if (!strcmp(k, "nranks")) nranks = atoi(v);
    else if (!strcmp(k, "read1")) read1 = v;
    else if (!strcmp(k, "read2")) read2 = v;
    else if (!strcmp(k, "diag_every")) diag_every = atoi(v);
    else if (!strcmp(k, "out")) out = v;
    else if (!strcmp(k, "device")) device = atoi(v);
    else if (!strcmp(k, "ic_only")) ic_only = atoi(v);
    else if (!strcmp(k, "mirror")) mirror = atoi(v);
    else if (strcmp(k, "model")) {
      fprintf(stderr, "chns2d: unknown parameter %s\n", k);
      return 2;
    }
  }
  const int Nx = p.Nx, Ny = p.Ny, nph = model == CHNS_MODEL_TERNARY ? 2 : (model == CHNS_MODEL_BINARY ? 1 : 0);
  const double dx = p.lx / Nx, dy = p.ly / Ny;
  if (p.epsilon <= 0.0) p.epsilon = 3.0 * dx;  /* global.h:103,131 */
  const int navg = (int)(maxiter / nfile);     /* main.c:52 */
  if (navg < 2 || diag_every < 1) {
    fprintf(stderr, "chns2d: need maxiter/nfile >= 2 and diag_every >= 1\n");
    return 2;
  }
  if (chdir(out) != 0) {
    fprintf(stderr, "chns2d: cannot enter %s\n", out);
    return 2;
  }
  mkdir("spectra", 0755);                      /* main.c:27-33 */
  mkdir("vort", 0755);
  mkdir("data", 0755);
  mkdir("init_field", 0755);
  if (nph) mkdir("phase_field", 0755);

  /* ---- initial condition (init.c) */
  const size_t npt = (size_t)Nx * Ny;
  double *f1 = nph ? malloc(npt * sizeof(double)) : NULL;
  double *f2 = nph == 2 ? malloc(npt * sizeof(double)) : NULL;
  if (!strcmp(init, "read")) {                 /* init.c:220-225, 252-262 */
    if (nph >= 1 && read_raw(read1, f1, npt)) { fprintf(stderr, "chns2d: cannot read %s\n", read1); return 1; }
    if (nph == 2 && read_raw(read2, f2, npt)) { fprintf(stderr, "chns2d: cannot read %s\n", read2); return 1; }
  } else if (!strcmp(init, "noise") && nph == 2) {   /* init.c:265-287, one slab per rank */
    if (Nx % nranks) { fprintf(stderr, "chns2d: nranks must divide Nx\n"); return 2; }
    const int ln = Nx / nranks;
    for (int r = 0; r < nranks; ++r) {
      srand(10 + r * 2);
      for (int i = r * ln; i < (r + 1) * ln; ++i)
        for (int j = 0; j < Ny; ++j) {
          const double rn1 = (double)rand() / (double)(RAND_MAX);
          const double rn2 = (double)rand() / (double)(RAND_MAX);
          f1[(size_t)i * Ny + j] = (1.0 / 4.0) + 0.001 * rn1;
          f2[(size_t)i * Ny + j] = (1.0 / 4.0) + 0.001 * rn2;
        }
    }
  } else if (!strcmp(init, "droplet") && nph == 1) {  /* init.c:227-243 */
    for (int i = 0; i < Nx; ++i)
      for (int j = 0; j < Ny; ++j) {
        const double x = i * dx, y = j * dy;
        const double r = sqrt((x - p.lx / 2.0) * (x - p.lx / 2.0) + (y - p.ly / 2.0) * (y - p.ly / 2.0));
        f1[(size_t)i * Ny + j] = 0.5 * (1.0 - tanh((r - (M_PI / 4.5)) / (0.5 * p.epsilon)));
      }
  } else if (strcmp(init, "zero") || nph) {
    fprintf(stderr, "chns2d: init=%s does not apply to this model\n", init);
    return 2;
  }
  if (ic_only) {
    if (nph >= 1) write_raw("init_field/ic1.bin", f1, npt);
    if (nph == 2) write_raw("init_field/ic2.bin", f2, npt);
    return 0;
  }
  /* ---- main.c:37-194 and init.c:291-294 through the reference-named entry
   * points of chns_ref.h (the state lives on the device; with mirror=1 the
   * host arrays cwz, cphi1, cphi2 are the state, as in the reference) */
  if (chns_ref_init(&p, device, mirror) != CHNS_OK) {
    fprintf(stderr, "chns2d: chns_ref_init failed (invalid parameters or no device)\n");
    return 1;
  }
  const size_t nc = (size_t)Nx * (Ny / 2 + 1);
  double (*cwz)[2] = malloc(nc * sizeof *cwz);
  double (*cphi1)[2] = nph >= 1 ? malloc(nc * sizeof *cphi1) : NULL;
  double (*cphi2)[2] = nph == 2 ? malloc(nc * sizeof *cphi2) : NULL;
  chns_ref_init_real(NULL, f1, f2, cwz, cphi1, cphi2);   /* zero vorticity (ZERO_INIT, init.c:89-100) */

  FILE *fu = fopen("data/ken_t.dat", "w");
  FILE *fe = nph ? fopen("data/free_en_t.dat", "w") : NULL;
  FILE *fdef1 = nph ? fopen(nph == 2 ? "data/gamma1.dat" : "data/gamma.dat", "w") : NULL;
  FILE *fdef2 = nph == 2 ? fopen("data/gamma2.dat", "w") : NULL;
  FILE *fph = nph ? fopen(nph == 2 ? "data/volume_phase1.dat" : "data/volume_droplet.dat", "w") : NULL;
  chns_ref_outputs(fu, fe, fdef1, fdef2);
  double t = 0.0;

  /* main.c:276-279 (every diag_every steps here) */
  #define DIAGNOSTICS()                          \
    do {                                         \
      chns_ref_set_time(t);                      \
      if (nph) free_energy(cphi1, cphi2);        \
      kinetic_energy(cwz, cphi1, cphi2);         \
      if (nph) perimeter(cphi1, cphi2);          \
    } while (0)

  DIAGNOSTICS();                               /* main.c:214-228 */
  spectrum(cwz, cphi1, cphi2);
  chns_ref_dump(0);                            /* main.c:230-250 (no truncation of the copy) */
  printf("\nTime stepping start.............\n");
  for (long it = 1; it <= (long)maxiter; ++it) {   /* main.c:260-367 */
    t += p.dt;
    chns_ref_set_time(t);
    etd2rk(cwz, cphi1, cphi2);
    if (it % diag_every == 0) DIAGNOSTICS();
    if (it % (navg / 2) == 0) spectrum(cwz, cphi1, cphi2);
    if (it % navg == 0) chns_ref_dump(1);
  }
  fclose(fu);
  if (fe) fclose(fe);
  if (fdef1) fclose(fdef1);
  if (fdef2) fclose(fdef2);
  if (fph) fclose(fph);
  chns_ref_finalize();
  free(cwz);
  free(cphi1);
  free(cphi2);
  free(f1);
  free(f2);
  return 0;
}

/*
 * chns2d - time-loop driver of the MI355X CHNS-2D step (C, on the C ABI).
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

static void die(const char *what, chns_ctx *c) {
  fprintf(stderr, "chns2d: %s: %s\n", what, c ? chns_last_error(c) : "");
  exit(1);
}

#define CK(call, ctx) \
  do {                \
    if ((call) != CHNS_OK) die(#call, ctx); \
  } while (0)

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
  int nranks = 1, diag_every = 1, device = 0, ic_only = 0;
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
  double *work = malloc(npt * sizeof(double));
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
  chns_ctx *ctx = NULL;
  if (chns_create(&ctx, &p, device) != CHNS_OK) die("chns_create", NULL);
  CK(chns_set_real_state(ctx, NULL, f1, f2), ctx);   /* zero vorticity (ZERO_INIT, init.c:89-100) + truncation */

  FILE *fu = fopen("data/ken_t.dat", "w");
  FILE *fe = nph ? fopen("data/free_en_t.dat", "w") : NULL;
  FILE *fdef1 = nph ? fopen(nph == 2 ? "data/gamma1.dat" : "data/gamma.dat", "w") : NULL;
  FILE *fdef2 = nph == 2 ? fopen("data/gamma2.dat", "w") : NULL;
  FILE *fph = nph ? fopen(nph == 2 ? "data/volume_phase1.dat" : "data/volume_droplet.dat", "w") : NULL;
  int nshell = 0, nxy = 0;
  chns_diag_sizes(&p, &nshell, &nxy);
  double *sp[7];
  for (int q = 0; q < 7; ++q) sp[q] = malloc((size_t)(q < 5 ? nshell : nxy) * sizeof(double));
  const double factor = (2 * M_PI) / p.lx;
  const int kxm = (int)lround(sqrt(((Nx / 2.0) * factor) * ((Nx / 2.0) * factor)));   /* main.c:62 */
  double t = 0.0;
  int ifile = 0, fl = 0;

  #define DIAGNOSTICS()                                                                             \
    do {                                                                                            \
      double eu = 0, e1 = 0, e2 = 0;                                                                \
      if (nph) {                                                                                    \
        double ephi, egphi;                                                                         \
        CK(chns_free_energy(ctx, &ephi, &egphi), ctx);                                              \
        fprintf(fe, "%lf \t %.12lf \t %.12lf\n", t, ephi, egphi);                                   \
        if (chns_status(ctx, NULL) == CHNS_E_DIVERGED) {  /* free_energy.c:151-154 */              \
          fprintf(stderr, "chns2d: %s\n", chns_last_error(ctx));                                    \
          exit(3);                                                                                  \
        }                                                                                           \
      }                                                                                             \
      CK(chns_kinetic_energy(ctx, &eu, &e1, &e2), ctx);                                             \
      if (nph == 2) fprintf(fu, "%lf \t %.12lf \t %.12lf \t %.12lf\n", t, eu, e1, e2);              \
      else if (nph == 1) fprintf(fu, "%lf \t %.12lf \t %.12lf\n", t, eu, e1);                      \
      else fprintf(fu, "%lf \t %.12lf\n", t, eu);                                                    \
      if (nph) {                                                                                    \
        double cnt[4];                                                                              \
        CK(chns_perimeter(ctx, cnt), ctx);                                                          \
        for (int ph = 0; ph < nph; ++ph) {                                                          \
          const double Spt = cnt[2 * ph] * dx, Sp0 = 2.0 * sqrt(M_PI * cnt[2 * ph + 1] * dx * dy);  \
          fprintf(ph == 0 ? fdef1 : fdef2, "%g \t %g \t %g \t %g\n", t, Sp0, Spt, (Spt / Sp0) - 1.0); \
        }                                                                                           \
      }                                                                                             \
    } while (0)

  #define SPECTRUM()                                                                                \
    do {                                                                                            \
      char fn[128];                                                                                 \
      CK(chns_spectrum(ctx, 1, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6]), ctx);             \
      snprintf(fn, sizeof fn, "spectra/spectrum.%d", ifile);                                        \
      FILE *fk = fopen(fn, "w");                                                                    \
      for (int ik = 1; ik < nshell; ++ik) {                                                         \
        const double kn = ik * (2 * M_PI / p.lx);                                                   \
        if (nph == 2) fprintf(fk, "%.5e \t %.12e \t %.12e \t %.12e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik], sp[2][ik], sp[3][ik], sp[4][ik]); \
        else if (nph == 1) fprintf(fk, "%.5e \t %.12e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik], sp[2][ik]); \
        else fprintf(fk, "%.5e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik]);                     \
      }                                                                                             \
      fclose(fk);                                                                                   \
      snprintf(fn, sizeof fn, "spectra/Xspectrum.%d", ifile);                                       \
      fk = fopen(fn, "w");                                                                          \
      for (int b = 1; b < kxm && b < nxy; ++b) fprintf(fk, "%d \t %.12e\n", b, sp[5][b]);           \
      fclose(fk);                                                                                   \
      snprintf(fn, sizeof fn, "spectra/Yspectrum.%d", ifile);                                       \
      fk = fopen(fn, "w");                                                                          \
      for (int b = 1; b < kxm && b < nxy; ++b) fprintf(fk, "%d \t %.12e\n", b, sp[6][b]);           \
      fclose(fk);                                                                                   \
      ++ifile;                                                                                      \
    } while (0)

  #define DUMP(trunc)                                                                               \
    do {                                                                                            \
      char fn[128];                                                                                 \
      CK(chns_real_field(ctx, 0, trunc, work), ctx);                                                \
      snprintf(fn, sizeof fn, "vort/wz.%d", fl);                                                    \
      write_raw(fn, work, npt);                                                                     \
      for (int ph = 1; ph <= nph; ++ph) {                                                           \
        CK(chns_real_field(ctx, ph, trunc, work), ctx);                                             \
        if (nph == 2) snprintf(fn, sizeof fn, "phase_field/phi%d.%d", ph, fl);                      \
        else snprintf(fn, sizeof fn, "phase_field/phi.%d", fl);                                     \
        write_raw(fn, work, npt);                                                                   \
      }                                                                                             \
      ++fl;                                                                                         \
    } while (0)

  DIAGNOSTICS();                               /* main.c:214-228 */
  SPECTRUM();
  DUMP(0);                                     /* main.c:230-250 (no truncation of the copy) */
  printf("\nTime stepping start.............\n");
  for (long it = 1; it <= (long)maxiter; ++it) {   /* main.c:260-367 */
    t += p.dt;
    CK(chns_step(ctx, 1), ctx);
    if (it % diag_every == 0) DIAGNOSTICS();
    if (it % (navg / 2) == 0) SPECTRUM();
    if (it % navg == 0) DUMP(1);
  }
  fclose(fu);
  if (fe) fclose(fe);
  if (fdef1) fclose(fdef1);
  if (fdef2) fclose(fdef2);
  if (fph) fclose(fph);
  chns_destroy(ctx);
  return 0;
}

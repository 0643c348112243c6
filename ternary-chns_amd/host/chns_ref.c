/*
 * chns_ref.c - the reference's time-loop entry points (etd2rk, free_energy,
 * kinetic_energy, perimeter, spectrum) with their names and argument lists,
 * over the C ABI (libchns_ref.so; declarations and modes in chns_ref.h).
 *
 * The diagnostics compute on the device (chns_free_energy, ...) and write the
 * reference's lines and files with its printf formats:
 *   kinetic_energy.c:109,135,141   free_energy.c:149   perimeter.c:123,264-265
 *   spectrum.c:150-187             write_array.c:48-79 (raw float64 dumps)
 */
#include "chns_ref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static struct {
  chns_ctx *ctx;
  chns_params p;
  int nph, mirror, nshell, nxy, kxm;
  FILE *fu, *fe, *fdef1, *fdef2;
  double t;
  int ifile, fl;            /* spectrum and dump file counters (global.c: ifile, fl) */
  double *sp[7];            /* ek, ensk, sph1, sph2, sp12, ekx, eky */
  double *work;             /* one packed real [Nx][Ny] field */
} g;

/* MPI_Abort's stand-in: the library's message, then exit. */
static void fail(const char *what, int status) {
  fprintf(stderr, "chns_ref: %s: %s\n", what, g.ctx ? chns_last_error(g.ctx) : "no context (chns_ref_init)");
  exit(status);
}

#define CK(call)                              \
  do {                                        \
    if (!g.ctx) fail(#call, 1);               \
    if ((call) != CHNS_OK) fail(#call, 1);    \
  } while (0)

static const double *cd(double (*a)[2]) { return a ? &a[0][0] : NULL; }
static double *md(double (*a)[2]) { return a ? &a[0][0] : NULL; }

/* mirror mode: the host arrays are the state */
static void upload(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  if (g.mirror) CK(chns_set_state(g.ctx, cd(cwz), cd(cphi1), cd(cphi2)));
}
static void writeback(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  if (g.mirror) CK(chns_get_state(g.ctx, md(cwz), md(cphi1), md(cphi2)));
}

int chns_ref_init(const chns_params *p, int device, int mirror) {
  if (g.ctx) chns_ref_finalize();
  const int rc = chns_create(&g.ctx, p, device);
  if (rc != CHNS_OK) {
    g.ctx = NULL;
    return rc;
  }
  g.p = *p;
  g.nph = p->model == CHNS_MODEL_TERNARY ? 2 : (p->model == CHNS_MODEL_BINARY ? 1 : 0);
  g.mirror = mirror != 0;
  g.t = 0.0;
  g.ifile = g.fl = 0;
  chns_diag_sizes(p, &g.nshell, &g.nxy);
  for (int q = 0; q < 7; ++q) g.sp[q] = malloc((size_t)(q < 5 ? g.nshell : g.nxy) * sizeof(double));
  g.work = malloc((size_t)p->Nx * p->Ny * sizeof(double));
  const double factor = (2 * M_PI) / p->lx;
  g.kxm = (int)lround(sqrt(((p->Nx / 2.0) * factor) * ((p->Nx / 2.0) * factor)));   /* main.c:62 */
  return CHNS_OK;
}

void chns_ref_finalize(void) {
  if (g.ctx) chns_destroy(g.ctx);
  g.ctx = NULL;
  for (int q = 0; q < 7; ++q) {
    free(g.sp[q]);
    g.sp[q] = NULL;
  }
  free(g.work);
  g.work = NULL;
}

void chns_ref_outputs(FILE *fu, FILE *fe, FILE *fdef1, FILE *fdef2) {
  g.fu = fu;
  g.fe = fe;
  g.fdef1 = fdef1;
  g.fdef2 = fdef2;
}

void chns_ref_set_time(double t) { g.t = t; }

void chns_ref_push(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  CK(chns_set_state(g.ctx, cd(cwz), cd(cphi1), cd(cphi2)));
}

void chns_ref_pull(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  CK(chns_get_state(g.ctx, md(cwz), md(cphi1), md(cphi2)));
}

void chns_ref_init_real(const double *wz, const double *phi1, const double *phi2, double (*cwz)[2],
                        double (*cphi1)[2], double (*cphi2)[2]) {
  CK(chns_set_real_state(g.ctx, wz, phi1, phi2));
  if (cwz || cphi1 || cphi2) CK(chns_get_state(g.ctx, md(cwz), md(cphi1), md(cphi2)));
}

void etd2rk(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  upload(cwz, cphi1, cphi2);
  CK(chns_step(g.ctx, 1));
  writeback(cwz, cphi1, cphi2);
}

void free_energy(double (*cphi1)[2], double (*cphi2)[2]) {
  if (!g.nph) return;
  if (g.mirror) CK(chns_set_state(g.ctx, NULL, cd(cphi1), cd(cphi2)));
  double ephi, egphi;
  CK(chns_free_energy(g.ctx, &ephi, &egphi));
  if (g.fe) fprintf(g.fe, "%lf \t %.12lf \t %.12lf\n", g.t, ephi, egphi);   /* free_energy.c:149 */
  if (chns_status(g.ctx, NULL) == CHNS_E_DIVERGED) {                        /* free_energy.c:151-154 */
    if (g.fe) fflush(g.fe);
    fail("free_energy", 3);
  }
}

void kinetic_energy(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  upload(cwz, cphi1, cphi2);
  double eu = 0, e1 = 0, e2 = 0, n1 = 0;
  CK(chns_kinetic_energy_counts(g.ctx, &eu, &e1, &e2, &n1, NULL));
  if (!g.fu) return;
  if (g.nph == 2) fprintf(g.fu, "%lf \t %.12lf \t %.12lf \t %.12lf\n", g.t, eu, e1, e2);   /* :135 */
  else if (g.nph == 1) {
    if (n1 != 0) fprintf(g.fu, "%lf \t %.12lf \t %.12lf\n", g.t, eu, e1);                 /* :103-109 */
  }
  else fprintf(g.fu, "%lf \t %.12lf\n", g.t, eu);                                          /* :141 */
}

void perimeter(double (*cphi1)[2], double (*cphi2)[2]) {
  if (!g.nph) return;
  if (g.mirror) CK(chns_set_state(g.ctx, NULL, cd(cphi1), cd(cphi2)));
  double cnt[4];
  CK(chns_perimeter(g.ctx, cnt));
  const double dx = g.p.lx / g.p.Nx, dy = g.p.ly / g.p.Ny;
  for (int ph = 0; ph < g.nph; ++ph) {
    FILE *f = ph == 0 ? g.fdef1 : g.fdef2;
    const double Spt = cnt[2 * ph] * dx, Sp0 = 2.0 * sqrt(M_PI * cnt[2 * ph + 1] * dx * dy);   /* :246-256 */
    if (f) fprintf(f, "%g \t %g \t %g \t %g\n", g.t, Sp0, Spt, (Spt / Sp0) - 1.0);            /* :264-265 */
  }
}

void spectrum(double (*cwz)[2], double (*cphi1)[2], double (*cphi2)[2]) {
  upload(cwz, cphi1, cphi2);
  double **sp = g.sp;
  CK(chns_spectrum(g.ctx, 1, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6]));
  char fn[128];
  snprintf(fn, sizeof fn, "spectra/spectrum.%d", g.ifile);   /* spectrum.c:150-151 */
  FILE *fk = fopen(fn, "w");
  if (!fk) fail("spectrum: cannot write spectra/", 1);
  for (int ik = 1; ik < g.nshell; ++ik) {
    const double kn = ik * (2 * M_PI / g.p.lx);
    if (g.nph == 2)
      fprintf(fk, "%.5e \t %.12e \t %.12e \t %.12e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik], sp[2][ik], sp[3][ik],
              sp[4][ik]);
    else if (g.nph == 1) fprintf(fk, "%.5e \t %.12e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik], sp[2][ik]);
    else fprintf(fk, "%.5e \t %.12e \t %.12e\n", kn, sp[0][ik], sp[1][ik]);
  }
  fclose(fk);
  /* spectrum.c:167-184: kxm - 1 lines each (the arrays hold N/2+2 entries) */
  for (int xy = 0; xy < 2; ++xy) {
    snprintf(fn, sizeof fn, xy == 0 ? "spectra/Xspectrum.%d" : "spectra/Yspectrum.%d", g.ifile);
    fk = fopen(fn, "w");
    if (!fk) fail("spectrum: cannot write spectra/", 1);
    for (int b = 1; b < g.kxm && b < g.nxy; ++b) fprintf(fk, "%d \t %.12e\n", b, sp[5 + xy][b]);
    fclose(fk);
  }
  ++g.ifile;
  writeback(cwz, cphi1, cphi2);   /* the truncated phase fields (spectrum.c:22-27) */
}

void chns_ref_dump(int truncate) {
  const size_t npt = (size_t)g.p.Nx * g.p.Ny;
  char fn[128];
  for (int which = 0; which <= g.nph; ++which) {
    CK(chns_real_field(g.ctx, which, truncate, g.work));
    if (which == 0) snprintf(fn, sizeof fn, "vort/wz.%d", g.fl);
    else if (g.nph == 2) snprintf(fn, sizeof fn, "phase_field/phi%d.%d", which, g.fl);
    else snprintf(fn, sizeof fn, "phase_field/phi.%d", g.fl);
    FILE *f = fopen(fn, "wb");
    if (!f) fail("chns_ref_dump: cannot write the dump", 1);
    if (fwrite(g.work, sizeof(double), npt, f) != npt) fail("chns_ref_dump: short write", 1);
    fclose(f);
  }
  ++g.fl;
}

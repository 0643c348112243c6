/*
 * chns_step.c - CPU restatement of the reference ETD2RK step in C with
 * OpenMP (TEST INFRASTRUCTURE: the checker for the HIP library at grid sizes
 * the numpy oracle is too slow for, and the CPU baseline of bench.py; the
 * product never links or calls it).
 *
 * It restates the same reference functions as oracle/chns_oracle.py, in the
 * reference's structure and FP64 arithmetic, with its own FFT instead of
 * FFTW (the reference's dependency, absent here):
 *   kxky            functions.c:11-32       truncation     functions.c:47-83 (N/4 circle)
 *   dfft / ifft     fft.c:11-31             get_velocity   velocity.c:3-44
 *   potential       chem_potential.c:12-94  surface_tension surface_tension.c:13-227
 *   nonlin          nonlin.c:13-381         etd2rk         etd2rk.c:14-224
 * Models NS, BINARY and TERNARY with the Kolmogorov forcing term of the
 * library (the reference's forcing.c does not compile); the optional physics
 * switches (gravity, walls, sponge, smooth filter, other dealias rules) live
 * in the numpy oracle only.  Two real fields share one complex transform
 * (a + ib), and the surface-tension force is summed in real space before its
 * transforms, so results equal the numpy oracle's to round-off, not bit for
 * bit (tests/test_c_oracle.py).
 *
 * Layouts are the reference's: complex half-spectra IDXC [N][N/2+1] (x slow,
 * global.h:13), real fields packed [N][N].  Square grids, N a power of two.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

enum { M_NS = 0, M_BINARY = 1, M_TERNARY = 2 };

/* ETD2RK coefficients (chns_tables_ref.c, compiled without FP contraction) */
void chnso_etd_mode(int model, double nu, double dt, double lambda, double epsilon, double sigma, double ksqr,
                    double *eta, double *z1w, double *z2w, double *beta, double *z1c, double *z2c);

typedef struct {
  int N, H, L, model, f_k, nph;
  double lx, nu, dt, sigma, eps, lam, s12, s13, s23, Xl, f_amp;
  double g1, g2, g3, del;
  cplx *tw;              /* exp(-2 pi i k/N), k < N/2 */
  int *rev;              /* bit reversal of log2 N bits */
  double *kx, *ky;
  unsigned char *keep;   /* truncation mask [N][H] */
  double *ks;            /* ksqr [N][H] */
  double *eta, *z1w, *z2w, *beta, *z1c, *z2c;   /* [N][H] */
  cplx *W, *C1, *C2;     /* state */
  cplx *S[12];           /* spectral scratch [N][H] */
  double *R[12];         /* real scratch [N][N] */
  cplx *T[2];            /* column-transform intermediates [N][H] */
} chnso;

/* ------------------------------------------------------------------ FFT */

/* In-place radix-2 FFT of length N (sign -1 forward, +1 inverse), unscaled. */
static void fft1(const chnso *o, cplx *a, int inverse) {
  const int N = o->N;
  for (int i = 0; i < N; ++i) {
    const int r = o->rev[i];
    if (r > i) {
      cplx t = a[i];
      a[i] = a[r];
      a[r] = t;
    }
  }
  for (int len = 2; len <= N; len <<= 1) {
    const int half = len >> 1, step = N / len;
    for (int s = 0; s < N; s += len) {
      for (int k = 0; k < half; ++k) {
        cplx w = o->tw[k * step];
        if (inverse) w = conj(w);
        const cplx u = a[s + k], v = a[s + k + half] * w;
        a[s + k] = u + v;
        a[s + k + half] = u - v;
      }
    }
  }
}

/* Column (x) transforms of the H columns of X [N][H] into Y (may alias X);
 * mask: apply truncation() to the input (functions.c:75-78). */
static void col_fft(const chnso *o, const cplx *X, cplx *Y, int inverse, int mask) {
  const int N = o->N, H = o->H;
#pragma omp parallel
  {
    cplx *buf = malloc((size_t)N * sizeof(cplx));
#pragma omp for schedule(static)
    for (int j = 0; j < H; ++j) {
      for (int i = 0; i < N; ++i) {
        const size_t e = (size_t)i * H + j;
        buf[i] = (!mask || o->keep[e]) ? X[e] : 0.0;
      }
      fft1(o, buf, inverse);
      for (int i = 0; i < N; ++i) Y[(size_t)i * H + j] = buf[i];
    }
    free(buf);
  }
}

/* ifft (fft.c:16-31) of truncation(A), truncation(B) (B may be NULL) into
 * real a, b: column iFFTs, then one complex row transform per row pair
 * (c2r: the imaginary parts at ky = 0 and ky = N/2 are ignored), scaled by
 * 1/(Nx*Ny). */
static void ifft2(chnso *o, const cplx *A, const cplx *B, double *a, double *b, int mask) {
  const int N = o->N, H = o->H;
  col_fft(o, A, o->T[0], 1, mask);
  if (B) col_fft(o, B, o->T[1], 1, mask);
  const double scale = 1.0 / ((double)N * N);
#pragma omp parallel
  {
    cplx *z = malloc((size_t)N * sizeof(cplx));
#pragma omp for schedule(static)
    for (int i = 0; i < N; ++i) {
      const cplx *x = o->T[0] + (size_t)i * H;
      const cplx *y = B ? o->T[1] + (size_t)i * H : NULL;
      for (int k = 0; k <= N / 2; ++k) {
        cplx xk = x[k], yk = y ? y[k] : 0.0;
        if (k == 0 || k == N / 2) {
          xk = creal(xk);
          yk = creal(yk);
        }
        z[k] = xk + I * yk;
        if (k > 0 && k < N / 2) z[N - k] = conj(xk) + I * conj(yk);
      }
      fft1(o, z, 1);
      for (int n = 0; n < N; ++n) {
        a[(size_t)i * N + n] = creal(z[n]) * scale;
        if (b) b[(size_t)i * N + n] = cimag(z[n]) * scale;
      }
    }
    free(z);
  }
}

/* dfft (fft.c:11-14) of real a, b (b may be NULL) into A, B: one complex row
 * transform of a + ib per row, split into the two half-spectra, then the
 * column transforms. */
static void dfft2(chnso *o, const double *a, const double *b, cplx *A, cplx *B) {
  const int N = o->N, H = o->H;
#pragma omp parallel
  {
    cplx *z = malloc((size_t)N * sizeof(cplx));
#pragma omp for schedule(static)
    for (int i = 0; i < N; ++i) {
      for (int n = 0; n < N; ++n) z[n] = a[(size_t)i * N + n] + I * (b ? b[(size_t)i * N + n] : 0.0);
      fft1(o, z, 0);
      for (int k = 0; k <= N / 2; ++k) {
        const cplx zk = z[k], zr = conj(z[(N - k) & (N - 1)]);
        A[(size_t)i * H + k] = 0.5 * (zk + zr);
        if (B) B[(size_t)i * H + k] = -0.5 * I * (zk - zr);
      }
    }
    free(z);
  }
  col_fft(o, A, A, 0, 0);
  if (B) col_fft(o, B, B, 0, 0);
}

/* ------------------------------------------------------------------ physics */

static size_t nc(const chnso *o) { return (size_t)o->N * o->H; }
static size_t nr(const chnso *o) { return (size_t)o->N * o->N; }

/* (I*k)*z with C99 semantics (real -k Im z, imaginary k Re z) */
static inline cplx ik(double k, cplx z) { return CMPLX(-k * cimag(z), k * creal(z)); }

/* The right-hand side N(u) of one stage: potential() then nonlin()
 * (etd2rk.c:40,51,149-150). */
static void rhs(chnso *o, const cplx *W, const cplx *C1, const cplx *C2, cplx *NW, cplx *NC1, cplx *NC2) {
  const int N = o->N, H = o->H;
  const size_t n = nc(o), m = nr(o);
  cplx *U = o->S[0], *V = o->S[1], *WX = o->S[2], *WY = o->S[3];
  double *ux = o->R[0], *uy = o->R[1], *wx = o->R[2], *wy = o->R[3];
  /* velocity (velocity.c:20-39) and grad w (nonlin.c:36-37) */
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) {
    const double kxv = o->kx[e / H], kyv = o->ky[e % H], ksv = o->ks[e];
    const cplx psi = ksv > 1e-12 ? CMPLX(creal(W[e]) / ksv, cimag(W[e]) / ksv) : 0.0;
    U[e] = ik(kyv, psi);
    V[e] = ik(-kxv, psi);
    WX[e] = ik(kxv, W[e]);
    WY[e] = ik(kyv, W[e]);
  }
  ifft2(o, U, V, ux, uy, 1);          /* nonlin.c:44-48 */
  ifft2(o, WX, WY, wx, wy, 1);
  double *nl = o->R[4];
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < m; ++e) nl[e] = ux[e] * wx[e] + uy[e] * wy[e];   /* nonlin.c:62 */
  dfft2(o, nl, NULL, NW, NULL);
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) NW[e] = -NW[e];                          /* nonlin.c:74,92 */

  if (o->model != M_NS) {
    const double eps = o->eps;
    cplx *FX = o->S[4], *FY = o->S[5];
    double *c1x = o->R[5], *c1y = o->R[6], *fx = o->R[7], *fy = o->R[8];
    cplx *A1 = o->S[6], *A2 = o->S[7];
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < n; ++e) {
      A1[e] = ik(o->kx[e / H], C1[e]);
      A2[e] = ik(o->ky[e % H], C1[e]);
    }
    ifft2(o, A1, A2, c1x, c1y, 1);       /* grad c1 (surface_tension.c:126-136, nonlin.c:272-284) */
    if (o->model == M_TERNARY) {
      const double g1 = o->g1, g2 = o->g2, g3 = o->g3;
      const double g1i = 1.0 / g1, g2i = 1.0 / g2, g3i = 1.0 / g3;   /* chem_potential.c:21-23 */
      double *c1 = o->R[9], *c2 = o->R[10];
      cplx *M1 = o->S[8], *M2 = o->S[9];
      /* potential(): muF from the truncated fields, transformed untruncated
       * (chem_potential.c:27-91) */
      ifft2(o, C1, C2, c1, c2, 1);
      const double X6 = 6.0 * o->Xl, fac = 2.0 * o->del / eps;
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < m; ++e) {
        const double a = c1[e], b = c2[e], c = (1.0 - a - b);
        const double abc = a * b * c;
        const double f1 = a * (1.0 - a) * (1.0 - 2.0 * a);
        const double f2 = b * (1.0 - b) * (1.0 - 2.0 * b);
        const double f3 = c * (1.0 - c) * (1.0 - 2.0 * c);
        const double G1 = g1 * f1 + X6 * abc * b * c;
        const double G2 = g2 * f2 + X6 * abc * a * c;
        const double G3 = g3 * f3 + X6 * abc * a * b;
        c1[e] = ((G1 - G2) * g2i + (G1 - G3) * g3i) * fac;   /* chem_potential.c:74-86 */
        c2[e] = ((G2 - G1) * g1i + (G2 - G3) * g3i) * fac;
      }
      dfft2(o, c1, c2, M1, M2);
      /* mu_i = T(M_i + 0.75 g_i eps k^2 c_i) (surface_tension.c:116-136) */
      cplx *MU1 = o->S[10], *MU2 = o->S[11];
      const double m1 = 0.75 * g1 * eps, m2 = 0.75 * g2 * eps;
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < n; ++e) {
        MU1[e] = M1[e] + C1[e] * (m1 * o->ks[e]);
        MU2[e] = M2[e] + C2[e] * (m2 * o->ks[e]);
      }
      double *mu1 = o->R[9], *mu2 = o->R[10];
      ifft2(o, MU1, MU2, mu1, mu2, 1);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < n; ++e) {
        A1[e] = ik(o->kx[e / H], C2[e]);
        A2[e] = ik(o->ky[e % H], C2[e]);
      }
      double *c2x = o->R[11], *c2y = o->R[4];
      ifft2(o, A1, A2, c2x, c2y, 1);     /* grad c2 (surface_tension.c:178-187, nonlin.c:314-326) */
      const double A11 = (1.0 + (g3 / g1)), A12 = (g3 / g2), A21 = (g3 / g1), A22 = (1.0 + (g3 / g2));
      double *ad1 = o->R[0], *ad2 = o->R[1];   /* overwrite ux, uy once used */
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < m; ++e) {
        const double a1 = A11 * mu1[e] + A12 * mu2[e];        /* surface_tension.c:144-145 */
        const double a2 = A21 * mu1[e] + A22 * mu2[e];        /* surface_tension.c:195-196 */
        fx[e] = c1x[e] * a1 + c2x[e] * a2;
        fy[e] = c1y[e] * a1 + c2y[e] * a2;
        const double u = ux[e], v = uy[e];
        ad1[e] = c1x[e] * u + c1y[e] * v;                      /* nonlin.c:292-300 */
        ad2[e] = c2x[e] * u + c2y[e] * v;                      /* nonlin.c:334-348 */
      }
      dfft2(o, fx, fy, FX, FY);
      dfft2(o, ad1, ad2, NC1, NC2);
      const double lg1 = o->lam / g1, lg2 = o->lam / g2;      /* nonlin.c:369-370 */
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < n; ++e) {
        NC1[e] = -NC1[e] - M1[e] * (lg1 * o->ks[e]);
        NC2[e] = -NC2[e] - M2[e] * (lg2 * o->ks[e]);
      }
    } else {
      /* binary: mu = T(1.5 sigma eps k^2 c + 24 sigma/eps F(P(Tc)))
       * (surface_tension.c:19-46), P(c) = c(1-c)(1-2c) */
      double *c = o->R[9];
      cplx *P = o->S[8], *MU = o->S[9];
      ifft2(o, C1, NULL, c, NULL, 1);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < m; ++e) c[e] = c[e] * (1.0 - c[e]) * (1.0 - 2.0 * c[e]);
      dfft2(o, c, NULL, P, NULL);
      const double mcb = 1.5 * o->sigma * eps, s24 = 24.0 * (o->sigma / eps);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < n; ++e) MU[e] = C1[e] * (mcb * o->ks[e]) + P[e] * s24;
      double *mu = o->R[10];
      ifft2(o, MU, NULL, mu, NULL, 1);
      double *ad = o->R[11];
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < m; ++e) {
        fx[e] = c1x[e] * mu[e];                                /* surface_tension.c:83-89 */
        fy[e] = c1y[e] * mu[e];
        ad[e] = ux[e] * c1x[e] + uy[e] * c1y[e];               /* nonlin.c:207 */
      }
      dfft2(o, fx, fy, FX, FY);
      dfft2(o, ad, NULL, NC1, NULL);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < n; ++e)
        NC1[e] = -NC1[e] - (P[e] * (s24 * o->ks[e])) * o->lam;   /* nonlin.c:240,251 */
    }
    /* curl of the force (curl.c:16-31): i (kx Fy - ky Fx), added to cnlz
     * (nonlin.c:104) */
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < n; ++e) NW[e] += I * (o->kx[e / H] * FY[e] - o->ky[e % H] * FX[e]);
  }
  if (o->f_amp != 0.0) NW[o->f_k] += CMPLX(0.0, -0.5 * o->f_amp * (double)N * (double)N);   /* (kx, ky) = (0, f_k) */
}

/* one ETD2RK step (etd2rk.c:14-224) */
static void step(chnso *o) {
  const size_t n = nc(o);
  cplx *n1[3], *a[3], *n2[3];
  /* stage buffers live past the rhs scratch: allocated once per call */
  const int nf = 1 + o->nph;
  cplx *buf = malloc(3 * nf * n * sizeof(cplx));
  for (int f = 0; f < nf; ++f) {
    n1[f] = buf + (size_t)f * n;
    a[f] = buf + (size_t)(nf + f) * n;
    n2[f] = buf + (size_t)(2 * nf + f) * n;
  }
  cplx *u[3] = {o->W, o->C1, o->C2};
  rhs(o, o->W, o->C1, o->C2, n1[0], nf > 1 ? n1[1] : NULL, nf > 2 ? n1[2] : NULL);
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) {
    a[0][e] = u[0][e] * o->eta[e] + n1[0][e] * o->z1w[e];                        /* :99 */
    for (int f = 1; f < nf; ++f) a[f][e] = u[f][e] * o->beta[e] + n1[f][e] * o->z1c[e];   /* :113,129-130 */
  }
  rhs(o, a[0], nf > 1 ? a[1] : NULL, nf > 2 ? a[2] : NULL, n2[0], nf > 1 ? n2[1] : NULL, nf > 2 ? n2[2] : NULL);
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < n; ++e) {
    u[0][e] = a[0][e] + (n2[0][e] - n1[0][e]) * o->z2w[e];                      /* :174 */
    for (int f = 1; f < nf; ++f) u[f][e] = a[f][e] + (n2[f][e] - n1[f][e]) * o->z2c[e];   /* :188,205-206 */
  }
  free(buf);
}

/* ------------------------------------------------------------------ C API (ctypes) */

/* dparams: lx, nu, dt, sigma, epsilon, lambda, sigma_12, sigma_13, sigma_23,
 * Xlambda, f_amp; iparams: N, model, f_k.  Returns NULL for an unsupported
 * grid. */
void *chnso_create(const double *dp, const int *ip) {
  const int N = ip[0];
  int L = 0;
  while ((1 << L) < N) ++L;
  if ((1 << L) != N || N < 8) return NULL;
  chnso *o = calloc(1, sizeof(chnso));
  o->N = N;
  o->H = N / 2 + 1;
  o->L = L;
  o->model = ip[1];
  o->f_k = ip[2];
  o->nph = o->model == M_TERNARY ? 2 : (o->model == M_BINARY ? 1 : 0);
  o->lx = dp[0]; o->nu = dp[1]; o->dt = dp[2]; o->sigma = dp[3]; o->eps = dp[4]; o->lam = dp[5];
  o->s12 = dp[6]; o->s13 = dp[7]; o->s23 = dp[8]; o->Xl = dp[9]; o->f_amp = dp[10];
  /* main.c:166-170 */
  o->g1 = (o->s12 - o->s23 + o->s13);
  o->g2 = (o->s12 - o->s13 + o->s23);
  o->g3 = (o->s23 - o->s12 + o->s13);
  o->del = (6.0 * o->g1 * o->g2 * o->g3);
  o->del /= (o->g1 * o->g2 + o->g1 * o->g3 + o->g2 * o->g3);
  const int H = o->H;
  const size_t n = nc(o), m = nr(o);
  o->tw = malloc((size_t)(N / 2) * sizeof(cplx));
  for (int k = 0; k < N / 2; ++k) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
    o->tw[k] = CMPLX((double)cosl(ang), (double)sinl(ang));
  }
  o->rev = malloc((size_t)N * sizeof(int));
  for (int i = 0; i < N; ++i) {
    int r = 0;
    for (int b = 0; b < L; ++b) r |= ((i >> b) & 1) << (L - 1 - b);
    o->rev[i] = r;
  }
  /* kxky (functions.c:11-32): Nyquist x-mode positive */
  const double f = (2 * M_PI) / o->lx;
  o->kx = malloc((size_t)N * sizeof(double));
  o->ky = malloc((size_t)H * sizeof(double));
  for (int i = 0; i < N; ++i) o->kx[i] = i <= N / 2 ? f * i : f * (i - N);
  for (int j = 0; j < H; ++j) o->ky[j] = f * j;
  const double dax = (N / 4.0) * f, dealias = dax * dax;   /* functions.c:56-59 */
  o->keep = malloc(n);
  o->ks = malloc(n * sizeof(double));
  double *tab[6];
  for (int q = 0; q < 6; ++q) tab[q] = malloc(n * sizeof(double));
  o->eta = tab[0]; o->z1w = tab[1]; o->z2w = tab[2]; o->beta = tab[3]; o->z1c = tab[4]; o->z2c = tab[5];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < H; ++j) {
      const size_t e = (size_t)i * H + j;
      const double ksqr = o->kx[i] * o->kx[i] + o->ky[j] * o->ky[j];
      o->ks[e] = ksqr;
      o->keep[e] = !(ksqr > dealias);                        /* functions.c:75-78 */
      chnso_etd_mode(o->model, o->nu, o->dt, o->lam, o->eps, o->sigma, ksqr, &o->eta[e], &o->z1w[e], &o->z2w[e],
                     &o->beta[e], &o->z1c[e], &o->z2c[e]);
    }
  o->W = calloc(n, sizeof(cplx));
  o->C1 = calloc(n, sizeof(cplx));
  o->C2 = calloc(n, sizeof(cplx));
  for (int q = 0; q < 12; ++q) {
    o->S[q] = calloc(n, sizeof(cplx));
    o->R[q] = calloc(m, sizeof(double));
  }
  o->T[0] = calloc(n, sizeof(cplx));
  o->T[1] = calloc(n, sizeof(cplx));
  return o;
}

void chnso_destroy(void *h) {
  chnso *o = h;
  if (!o) return;
  free(o->tw); free(o->rev); free(o->kx); free(o->ky); free(o->keep); free(o->ks);
  free(o->eta); free(o->z1w); free(o->z2w); free(o->beta); free(o->z1c); free(o->z2c);
  free(o->W); free(o->C1); free(o->C2);
  for (int q = 0; q < 12; ++q) {
    free(o->S[q]);
    free(o->R[q]);
  }
  free(o->T[0]);
  free(o->T[1]);
  free(o);
}

/* state in/out: interleaved (re, im) IDXC arrays; phase fields may be NULL */
void chnso_set_state(void *h, const double *w, const double *c1, const double *c2) {
  chnso *o = h;
  const size_t bytes = nc(o) * sizeof(cplx);
  memcpy(o->W, w, bytes);
  if (c1 && o->nph >= 1) memcpy(o->C1, c1, bytes);
  if (c2 && o->nph >= 2) memcpy(o->C2, c2, bytes);
}

void chnso_get_state(void *h, double *w, double *c1, double *c2) {
  chnso *o = h;
  const size_t bytes = nc(o) * sizeof(cplx);
  memcpy(w, o->W, bytes);
  if (c1 && o->nph >= 1) memcpy(c1, o->C1, bytes);
  if (c2 && o->nph >= 2) memcpy(c2, o->C2, bytes);
}

void chnso_step(void *h, int nsteps) {
  for (int s = 0; s < nsteps; ++s) step((chnso *)h);
}

/* truncation of the phase fields in place (spectrum.c:22-27) */
void chnso_truncate_state(void *h) {
  chnso *o = h;
  const size_t n = nc(o);
  for (size_t e = 0; e < n; ++e)
    if (!o->keep[e]) {
      o->C1[e] = 0.0;
      o->C2[e] = 0.0;
    }
}

/* OpenMP threads for the following calls (0: leave as is); returns the count in use */
int chnso_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

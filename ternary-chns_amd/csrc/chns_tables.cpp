// Host-built tables and parameter handling (no HIP; compiled with
// -ffp-contract=off so every expression rounds exactly like the reference,
// which is built by mpicc with no flags: plain IEEE double, no FMA).
//
//  * ETD2RK coefficients: etd2rk.c:88-127 (stage 1) and :162-203 (stage 2).
//    The phi_2 coefficient cancels catastrophically at low k, so the values
//    are only reproducible with the reference's operation order and the C
//    library's exp(); the device never evaluates them.
//  * Dealias extents: truncation() predicate ksqr > dax*dax (functions.c:47-83).
//  * Twiddles exp(-2 pi i n/N), evaluated in long double and rounded once.
#include <cmath>
#include <cstring>
#include <vector>

#include "chns.h"
#include "chns_host.hpp"

extern "C" void chns_params_default(chns_params* p, int model) {
  std::memset(p, 0, sizeof(*p));
  p->Nx = 1024;               // global.h:23-24
  p->Ny = 1024;
  p->model = model;
  p->lx = 4.0 * M_PI;         // global.h:20-21
  p->ly = 4.0 * M_PI;
  p->nu = 0.01;               // global.h:30
  p->dt = 1e-4;               // global.h:31
  p->epsilon = -1.0;          // -> 3.0*dx (global.h:103,131)
  if (model == CHNS_MODEL_BINARY) {
    p->sigma = 2.0;           // global.h:102
    p->lambda = 1e-6;         // global.h:104
  } else {
    p->sigma = 2.0;
    p->lambda = 1e-4;         // global.h:132
  }
  p->sigma_12 = 0.2;          // global.h:127-129
  p->sigma_23 = 0.45;
  p->sigma_13 = 0.2;
  p->Xlambda = 2.0;           // global.h:133
  p->smooth_filter = 0;
  p->f_amp = 0.0;
  p->f_k = 0;
  p->gravity = 0;
  if (model == CHNS_MODEL_BINARY) {
    p->rho1 = 1.044;          // global.h:108-110
    p->rho2 = 1.0;
  } else {
    p->rho1 = 0.957;          // global.h:137-140
    p->rho2 = 1.044;
  }
  p->rho3 = 1.0;
  p->g = 3.0;
  p->vpm = 0;
  p->sponge = 0;
  p->dealias = 0;             // functions.c:56-59: the circular N/4 cut
}

namespace chns_host {

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

const char* check_params(const chns_params* p) {
  if (!p) return "null params";
  if (p->model < 0 || p->model > 2) return "model must be 0 (NS), 1 (BINARY) or 2 (TERNARY)";
  if (p->Nx != p->Ny) return "the library requires Nx == Ny";
  if (p->lx != p->ly) return "the library requires lx == ly";
  const int l = ilog2(p->Nx);
  if (l < 4 || l > 11) return "Nx must be a power of two in [16, 2048]";
  if (!(p->lx > 0.0)) return "lx must be positive";
  if (!(p->dt > 0.0)) return "dt must be positive";
  if (p->f_amp != 0.0 && !(p->f_k > 0 && p->f_k < p->Ny / 2)) return "forcing f_k must lie in (0, Ny/2)";
  if (p->gravity && p->model == CHNS_MODEL_NS) return "gravity needs a phase field (BINARY or TERNARY)";
  if (p->sponge && !p->vpm) return "sponge needs vpm (its profile is set up with the walls, init.c:184-211)";
  if ((p->vpm || p->sponge) && !(p->dt > 0.0)) return "vpm needs dt > 0 (eta = 1.5 dt)";
  if (p->dealias < 0 || p->dealias > 2) return "dealias must be 0 (N/4 circle), 1 (2/3 rule) or 2 (N/3 circle)";
  if (p->dealias != 0 && p->smooth_filter) return "smooth_filter replaces the cut: use dealias 0 with it";
  if (p->model == CHNS_MODEL_TERNARY) {
    Derived d;
    derive(p, &d);
    if (d.g1 == 0.0 || d.g2 == 0.0 || d.g3 == 0.0) return "ternary gammas must be non-zero";
  }
  return nullptr;
}

// chi(x_i) and sponge(y_j) (the latter times 1/eta) of init.c:184-211:
// eps = 3 dx, walls at 6 dx and l - 6 dx, eta = 1.5 dt (global.h:48).
void build_wall_tables(const chns_params* p, double* chi, double* sponge) {
  const double dx = p->lx / p->Nx, dy = p->ly / p->Ny;
  const double eps = 3.0 * dx;
  const double yd = 6.0 * dy, xd = 6.0 * dx;
  const double yu = p->ly - yd, xu = p->lx - xd;
  const double eta_inv = 1.0 / (1.5 * p->dt);
  for (int i = 0; i < p->Nx; ++i) {
    const double x = i * dx;
    chi[i] = 0.5 * (1.0 - tanh(2.0 * (x - xd) / eps)) + 0.5 * (1.0 + tanh(2.0 * (x - xu) / eps));
  }
  for (int j = 0; j < p->Ny; ++j) {
    const double y = j * dy;
    sponge[j] = 0.5 * (1.0 - tanh(2.0 * (y - yd) / eps)) + 0.5 * (1.0 + tanh(2.0 * (y - yu) / eps));
    sponge[j] *= eta_inv;
  }
}

void derive(const chns_params* p, Derived* d) {
  d->N = p->Nx;
  d->log2n = ilog2(p->Nx);
  d->H = p->Ny / 2 + 1;
  d->dx = p->lx / p->Nx;                               // global.h:27
  d->eps = p->epsilon > 0.0 ? p->epsilon : 3.0 * d->dx;  // global.h:103,131
  d->factor_x = (2 * M_PI) / p->lx;                    // main.c:54
  d->factor_y = (2 * M_PI) / p->ly;                    // main.c:55
  // main.c:166-170
  d->g1 = (p->sigma_12 - p->sigma_23 + p->sigma_13);
  d->g2 = (p->sigma_12 - p->sigma_13 + p->sigma_23);
  d->g3 = (p->sigma_23 - p->sigma_12 + p->sigma_13);
  d->del = (6.0 * d->g1 * d->g2 * d->g3);
  d->del /= (d->g1 * d->g2 + d->g1 * d->g3 + d->g2 * d->g3);
}

static inline double kx_of(const chns_params* p, const Derived& d, int i) {
  const int i1 = i;                                    // local_x_start = 0 on one device
  return (i1 <= p->Nx / 2) ? d.factor_x * i1 : d.factor_x * (i1 - p->Nx);   // functions.c:19-25
}
static inline double ky_of(const Derived& d, int j) { return d.factor_y * j; }   // functions.c:29-31

// One mode's coefficients, written exactly as etd2rk.c states them.
static void etd_mode(const chns_params* p, const Derived& d, double ksqr, double* eta_o, double* z1w_o,
                     double* z2w_o, double* beta_o, double* z1c_o, double* z2c_o) {
  const double nu = p->nu, dt = p->dt, lambda = p->lambda, epsilon = d.eps, sigma = p->sigma;
  double c, eta, zeta, beta;
  // stage 1, omega (etd2rk.c:90-97)
  c = -nu * ksqr;
  eta = exp(-nu * dt * ksqr);
  if (fabs(c) > 1e-12) zeta = (eta - 1.0) / c;
  else zeta = dt;
  *eta_o = eta;
  *z1w_o = zeta;
  // stage 2, omega (etd2rk.c:164-172)
  if (fabs(c) > 1e-12) zeta = (eta - 1.0 - c * dt) / (dt * c * c);
  else zeta = 0.5 * dt;
  *z2w_o = zeta;
  if (p->model == CHNS_MODEL_NS) return;
  if (p->model == CHNS_MODEL_BINARY) {
    // etd2rk.c:102-111 and :177-186 (stage 2 tests c != 0.0)
    c = -1.5 * lambda * sigma * epsilon * ksqr * ksqr;
    beta = exp(dt * c);
    if (fabs(c) > 1e-12) zeta = (beta - 1.0) / c;
    else zeta = dt;
    *beta_o = beta;
    *z1c_o = zeta;
    if (fabs(c) != 0.0) zeta = (beta - 1.0 - c * dt) / (dt * c * c);
    else zeta = 0.5 * dt;
    *z2c_o = zeta;
  } else {
    // etd2rk.c:117-127 and :191-203
    c = -0.75 * lambda * epsilon;
    c *= ksqr * ksqr;
    beta = exp(dt * c);
    if (fabs(c) > 1e-12) zeta = (beta - 1.0) / c;
    else zeta = dt;
    *beta_o = beta;
    *z1c_o = zeta;
    if (fabs(c) > 1e-12) {
      zeta = (beta - 1.0 - c * dt);
      zeta /= (dt * c * c);
    } else {
      zeta = 0.5 * dt;
    }
    *z2c_o = zeta;
  }
}

void build_coef_tables(const chns_params* p, double* tw1, double* tc1, double* tw2, double* tc2) {
  Derived d;
  derive(p, &d);
  const int H = d.H;
  for (int j = 0; j < H; ++j) {
    for (int ii = 0; ii < H; ++ii) {
      const double kx = kx_of(p, d, ii), ky = ky_of(d, j);
      const double ksqr = kx * kx + ky * ky;
      double e = 0, z1w = 0, z2w = 0, b = 0, z1c = 0, z2c = 0;
      etd_mode(p, d, ksqr, &e, &z1w, &z2w, &b, &z1c, &z2c);
      const size_t o = (size_t)j * H + ii;
      tw1[2 * o] = e; tw1[2 * o + 1] = z1w;
      tc1[2 * o] = b; tc1[2 * o + 1] = z1c;
      tw2[o] = z2w;
      tc2[o] = z2c;
    }
  }
}

void build_twiddles(int N, double* tw) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int n = 0; n < N; ++n) {
    const long double a = -2.0L * pi * (long double)n / (long double)N;
    tw[2 * n] = (double)cosl(a);
    tw[2 * n + 1] = (double)sinl(a);
  }
}

void dealias_extent(const chns_params* p, int* amax) {
  Derived d;
  derive(p, &d);
  if (p->smooth_filter) {
    // the smooth filter zeroes no mode (functions.c:70-73)
    for (int j = 0; j < d.H; ++j) amax[j] = p->Nx / 2;
    return;
  }
  if (p->dealias == 1) {
    // 2/3 rule: zero where |kx| > (Nx/3) 2pi/lx or |ky| > (Ny/3) 2pi/ly
    const double dax = (p->Nx / 3.0) * d.factor_x, day = (p->Ny / 3.0) * d.factor_y;
    for (int j = 0; j < d.H; ++j) {
      int a = -1;
      if (!(ky_of(d, j) > day))
        for (int i = 0; i <= p->Nx / 2; ++i)
          if (!(kx_of(p, d, i) > dax)) a = i;
      amax[j] = a;
    }
    return;
  }
  // the reference's circular cut (functions.c:56-59, radius Nx/4), or the
  // same predicate at radius Nx/3 (dealias 2)
  const double dax = (p->Nx / (p->dealias == 2 ? 3.0 : 4.0)) * d.factor_x;   // functions.c:56
  const double dealias = dax * dax;                    // functions.c:59
  for (int j = 0; j < d.H; ++j) {
    int a = -1;
    const double ky = ky_of(d, j);
    for (int i = 0; i <= p->Nx / 2; ++i) {
      const double kx = kx_of(p, d, i);
      const double ksqr = kx * kx + ky * ky;           // functions.c:69
      if (!(ksqr > dealias)) a = i;                    // functions.c:75-78
    }
    amax[j] = a;
  }
}

void build_cut_table(const chns_params* p, const int* amax, double* ft) {
  const int H = p->Ny / 2 + 1;
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < H; ++i) ft[(size_t)j * H + i] = i <= amax[j] ? 1.0 : 0.0;
}

void build_filter_table(const chns_params* p, double* ft) {
  Derived d;
  derive(p, &d);
  const double kxmax = (p->Nx / 2.0) * d.factor_x;             // main.c:57
  const double kymax = (p->Ny / 2.0) * d.factor_y;             // main.c:58
  const double ksqrmax = (double)(kxmax * kxmax + kymax * kymax);   // functions.c:54
  for (int j = 0; j < d.H; ++j) {
    const double ky = ky_of(d, j);
    for (int i = 0; i <= p->Nx / 2; ++i) {
      const double kx = kx_of(p, d, i);
      const double ksqr = kx * kx + ky * ky;                   // functions.c:69
      const double pp = ksqr / ksqrmax;                        // functions.c:71
      ft[(size_t)j * d.H + i] = exp(-36.0 * pow(pp, 18));      // functions.c:72-73
    }
  }
}

}  // namespace chns_host

extern "C" int chns_etd_tables(const chns_params* p, double* eta, double* z1w, double* z2w, double* beta,
                               double* z1c, double* z2c) {
  if (chns_host::check_params(p)) return CHNS_E_ARG;
  if (!eta || !z1w || !z2w) return CHNS_E_ARG;
  if (p->model != CHNS_MODEL_NS && (!beta || !z1c || !z2c)) return CHNS_E_ARG;
  chns_host::Derived d;
  chns_host::derive(p, &d);
  for (int i = 0; i < p->Nx; ++i) {
    for (int j = 0; j < d.H; ++j) {
      const double kx = chns_host::kx_of(p, d, i), ky = chns_host::ky_of(d, j);
      const double ksqr = kx * kx + ky * ky;           // etd2rk.c:88
      const size_t o = (size_t)i * d.H + j;            // IDXC
      double e, a, b, bb, c1, c2;
      chns_host::etd_mode(p, d, ksqr, &e, &a, &b, &bb, &c1, &c2);
      eta[o] = e; z1w[o] = a; z2w[o] = b;
      if (p->model != CHNS_MODEL_NS) { beta[o] = bb; z1c[o] = c1; z2c[o] = c2; }
    }
  }
  return CHNS_OK;
}

extern "C" int chns_filter_table(const chns_params* p, double* ft) {
  if (chns_host::check_params(p) || !ft) return CHNS_E_ARG;
  chns_host::Derived d;
  chns_host::derive(p, &d);
  std::vector<double> t((size_t)d.H * d.H);
  chns_host::build_filter_table(p, t.data());
  for (int i = 0; i < p->Nx; ++i)
    for (int j = 0; j < d.H; ++j) ft[(size_t)i * d.H + j] = t[(size_t)j * d.H + (i <= p->Nx / 2 ? i : p->Nx - i)];
  return CHNS_OK;
}

extern "C" int chns_dealias_extent(const chns_params* p, int* amax) {
  if (chns_host::check_params(p) || !amax) return CHNS_E_ARG;
  chns_host::dealias_extent(p, amax);
  return CHNS_OK;
}

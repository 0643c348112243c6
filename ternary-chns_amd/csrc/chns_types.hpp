// Plain-data types shared by the kernels (hipcc device pass) and the host API.
#pragma once

namespace chns {

struct cd { double x, y; };   // complex double, same layout as double[2]


struct KParams {
  int N;         // grid size, Nx = Ny = N
  int ldJ;       // leading dimension of row-pass inputs (>= Jt, multiple of 4)
  int ldF;       // leading dimension of row-pass outputs (>= N/2+1, multiple of 4)
  int Jt;        // number of columns holding kept (dealiased) modes
  int f_k;       // forced ky column (0: no forcing)
  double f;      // 2*pi/L
  double scale;  // 1/(Nx*Ny)
  double force_im;             // Im of the forcing half-spectrum at (0, f_k)
  // ternary (chem_potential.c:21-23,74-86; surface_tension.c:123-124,144-145,195-196; nonlin.c:369-370)
  double g1, g2, g3, g1i, g2i, g3i, fac, X6;
  double mc1, mc2;             // 0.75*gamma_i*epsilon
  double A11, A12, A21, A22;   // 1+g3/g1, g3/g2, g3/g1, 1+g3/g2
  double lg1, lg2;             // lambda/gamma_i
  // binary (surface_tension.c:141-142; nonlin.c:240,251)
  double mcb;                  // 1.5*sigma*epsilon
  double s24;                  // 24*(sigma/epsilon)
  double lam;
  // optional physics (macro.h GRAVITY, VPM, SPONGE)
  int grav, vpm, sponge;
  double gr1, gr2;             // -g (rho1 - rho3), -g (rho2 - rho3); binary -g (rho1 - rho2), 0
  double etainv;               // 1/eta, eta = 1.5 dt
};

struct Bufs {
  cd *W, *C1, *C2;             // state (column-major [Ny2][N])
  cd *Wo, *C1o, *C2o;          // where pass E writes the updated state: W, C1, C2 themselves
                               // except in the adjoint's trajectory replay
  cd *NW, *NC1, *NC2;          // stage 1's a - z2 N_n (N_n is saved at etd2rk.c:62-65)
  cd *Hc1, *Hc2;               // row-major [N][ldJ]
  cd *Hu, *Hv, *Hwx, *Hwy, *Hmu1, *Hmu2, *Hdx1, *Hdx2;
  cd *Mh1, *Mh2;               // row-major [N][ldF]
  cd *G1, *G2, *G3, *G4;       // W1, Fy, adv1, adv2 (row-major [N][ldF]); conservative pass D:
                               // -i ky F(Fx + v w), F(Fy - u w), adv1, adv2
  const cd* tw;                // twiddles exp(-2 pi i n/N)
  // ETD2RK coefficients, [(N/2+1)*(N/2+1)] indexed j*(N/2+1) + |i|, one
  // table per field kind so each field group reads only its own
  const cd* tabw1;             // omega stage 1: {eta, z1}
  const cd* tabc1;             // phase stage 1: {beta, z1}
  const double* tabw2;         // omega stage 2: z2
  const double* tabc2;         // phase stage 2: z2
  const int* amask;            // [Ny2] largest kept |i| per column, -1 if none
  const double* filt;          // [(N/2+1)*(N/2+1)] truncation factor at j*(N/2+1) + |i|, applied twice
                               // wherever truncation() is: SMOOTH_FILTER's exp(-36 pp^18), or 1/0 for
                               // the 2/3-rule and N/3 cuts (null: the pruned N/4 cut of the reference)
  const double* chi;           // [N] wall profile chi(x) (vpm)
  const double* spg;           // [N] sponge(y)/eta (sponge)
  cd* GR;                      // buoyancy torque, column-major (gravity)
  cd* Hw;                      // column iFFT of T(omega), TJ tiled (sponge; conservative pass D)
  int* flag;                   // divergence flag: set to 1 by pass E when an updated mode is not finite
};

// ---- diagnostics (chns_diag.hpp)
// per-mode operator of the full column iFFT
enum { COP_ID = 0, COP_IKX = 1, COP_IKY = 2, COP_U = 3, COP_V = 4 };
// reduction kinds (up to RED_Q sums per kind)
constexpr int RED_Q = 5;
enum { RED_FE_TER = 0, RED_FE_BIN = 1, RED_KE_TER = 2, RED_KE_BIN = 3, RED_KE_NS = 4, RED_PERIM = 5 };

struct RedArgs {
  const double* f[6];   // packed [N][N] fields (meaning per kind)
  double g1, g2, g3;    // gammas (ternary)
  int N;
};

// spectrum.c binning arguments
struct SpecArgs {
  const cd *W, *C1, *C2;   // column-major state
  const int* amask;
  const double* filt;      // SMOOTH_FILTER factor table (Bufs::filt) or null
  double f, dk, kymax, scale2;
  int N, nsh, nxy, nph;
  double* part;    // [5][nsh][N]
  double* party;   // [nxy][N]
  double* partx;   // [N]
};

}  // namespace chns

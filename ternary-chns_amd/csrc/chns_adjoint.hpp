// Discrete adjoint of the ternary ETD2RK step (FP64, gfx950).
//
// Forward step (etd2rk.c:14-224): a = E u + Z1 N(u), u' = a + Z2 (N(a) - N(u)).
// Reverse step, given ubar = dL/du':
//   abar = ubar + J_N(a)^T (Z2 ubar)
//   dL/du = E abar + J_N(u)^T (Z1 abar - Z2 ubar)
// A spectral adjoint Xbar is the unnormalised r2c of the real-space gradient
// dL/dx of the field x = c2r(X)/N^2, so the transpose of a real Fourier
// multiplier m(k) is conj(m(k)), of a c2r is an r2c (and vice versa) and
// pointwise products transpose pointwise.  Every read of the state in N goes
// through truncation(), so adjoint states live on the kept modes and every
// adjoint transform is pruned like the forward ones.
//
// One vector-Jacobian product J_N(x)^T g at a linearisation point x reuses
// the forward passes A, B, C at x (chns_kernels.hpp) for the real-space
// fields c_i, grad c_i, u, v, grad w, mu_i, then:
//   adj_rowD1: row c2r of xi = c2r(g_w), eta_i = c2r(g_i) and of u, v, grad w,
//              grad c_i; transposes of u.grad w and u.grad c_i; row r2c of
//              ubar, vbar, the grad-w bars and the advective grad-c bars
//   adj_rowD2: row c2r of d_x xi, d_y xi, grad c_i, mu_i; transpose of the
//              surface-tension force F = A1 grad c1 + A2 grad c2; row r2c of
//              its grad-c bars and of mubar_i
//   adj_colBC: column FFT of r2c(mubar_i), Mbar_i = T(.) - (lambda/g_i) k^2 g_i,
//              column iFFT -> Hmb_i
//   adj_rowB : row c2r of c_i and muFbar_i, transpose of the chemical potential
//              (chem_potential.c:40-86), row r2c of cbar_i -> Ocb_i
//   adj_colE : column FFTs, J^T g per field, the ETD2RK transpose, then the
//              column iFFTs of the next product's input (xi, d_x xi, eta_i).
// Coefficient partials for d/dsigma (sums over the grid) are written per
// workgroup into fixed slots and summed in a fixed order (k_adj_accum), so
// results are bitwise repeatable.
#pragma once
#include "chns_adjoint_types.hpp"
#include "chns_kernels.hpp"

namespace chns {

// Workgroup sum of s[0..NQ) into dst[q*stride] through the LDS array red
// (>= NQ*NT doubles; a fixed tree, so bitwise repeatable).  All NT threads call it.
template <int NQ, int NT>
__device__ __forceinline__ void block_sum(const double* s, double* red, double* dst, size_t stride) {
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) red[q * NT + tid] = s[q];
  __syncthreads();
#pragma unroll
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q * NT + tid] += red[q * NT + tid + w];
    }
    __syncthreads();
  }
  if (tid < NQ) dst[tid * stride] = red[tid * NT];
}

// Row-pass input loader shared by the adjoint row kernels: at one row per
// workgroup (N = 2048) each field pair is staged in LDS by the DMA row stager
// while the previous pair is transformed; otherwise pairs load directly.
// Pair k: X_k, Y_k (Y may be null or equal to X), YIK_k: Y multiplied by i ky.
template <int LOG2N>
struct AdjRowLoader {
  static constexpr bool STG = RowCfg<LOG2N>::R == 1;
  RowStager<LOG2N> st;
  size_t ro;
  int Jt;
  double f;
  // first pair: stage and assemble it, then start staging the second
  template <bool YIK, bool HASY, bool YIK2, bool HASY2>
  __device__ __forceinline__ void first(const cd* X, const cd* Y, const cd* X2, const cd* Y2, int t, cd* v) const {
    if constexpr (STG) {
      const bool same = (X == Y);
      if (same) st.template issue<false>(X + ro, nullptr);
      else st.template issue<HASY>(X + ro, HASY ? Y + ro : nullptr);
      st.wait();
      if (same) assemble_pair<LOG2N, YIK>(st.stg, st.stg, 1, Jt, f, t, v);
      else st.template assemble<YIK, HASY>(f, t, v);
      lds_sync();
      st.template issue<HASY2>(X2 + ro, HASY2 ? Y2 + ro : nullptr);
    } else {
      row_load_pair<LOG2N, YIK>(X + ro, HASY ? Y + ro : nullptr, Jt, f, t, v);
    }
  }
  // pair staged by the previous call (X, Y); start staging (X2, Y2) unless LAST
  template <bool YIK, bool HASY, bool HASY2, bool LAST>
  __device__ __forceinline__ void next(const cd* X, const cd* Y, const cd* X2, const cd* Y2, int t, cd* v) const {
    if constexpr (STG) {
      st.wait();
      st.template assemble<YIK, HASY>(f, t, v);
      if constexpr (!LAST) {
        lds_sync();
        st.template issue<HASY2>(X2 + ro, HASY2 ? Y2 + ro : nullptr);
      }
    } else {
      row_load_pair<LOG2N, YIK>(X + ro, HASY ? Y + ro : nullptr, Jt, f, t, v);
    }
  }
};

// ============================================================ adj_rowD1
// Transposes of the advection terms: N_w = -R(u w_x + v w_y) + ...,
// N_i = -R(u c_ix + v c_iy) - ...  with nlzbar = -xi, advbar_i = -eta_i
// (xi = c2r(g_w), eta_i = c2r(g_i)).  Outputs: O0 = -i ky R(ubar), O1 = R(vbar),
// O2 = R(w_x bar), O3 = -i ky R(w_y bar), and the advective parts of the grad-c
// bars: O10 = R(c1x bar), O11 = -i ky R(c1y bar), O12, O13 likewise for c2.
// grid: N / R
template <int LOG2N>
__global__ __launch_bounds__(RowCfg<LOG2N>::NT, 2) void k_adj_rowD1(AdjArgs args) {
  const AdjBufs& A = args.A;
  const KParams& P = args.P;
  using RC = RowCfg<LOG2N>;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  constexpr bool STG = AdjRowLoader<LOG2N>::STG;
  __shared__ cd stg_all[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;
  const int x = row_of<LOG2N>(r);
  cd* lds = lds_all + r;
  const size_t fo = tix<TF>(x, 0, P.ldJ);
  const double sc = P.scale;
  const int Jt = P.Jt;
  const AdjRowLoader<LOG2N> ld{{stg_all, Jt}, tix<TJ>(x, 0, P.ldJ), Jt, P.f};
  cd v[8];
  RowTw<LOG2N, RC::R> tws;
  tws.template init<STG>(A.tw, t);   // before the stager's first copy
  double xi[8], e1[8], e2[8];
  ld.template first<false, true, false, false>(A.Hxi, A.Het1, A.Het2, nullptr, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) { xi[m] = v[m].x * sc; e1[m] = v[m].y * sc; }
  ld.template next<false, false, true, false>(A.Het2, nullptr, A.Hu, A.Hv, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) e2[m] = v[m].x * sc;
  // (u, v): w bars and the advective grad-c bars, each pair to a row r2c right away
  double u[8], w[8];
  ld.template next<false, true, true, false>(A.Hu, A.Hv, A.Hwx, A.Hwy, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    u[m] = v[m].x * sc;
    w[m] = v[m].y * sc;
    v[m] = mk(-xi[m] * u[m], -xi[m] * w[m]);
  }
  {
    cd* o2 = A.O[2] + fo;
    cd* o3 = A.O[3] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
      if (n < Jt) {
        o2[TF * n] = a;
        o3[TF * n] = ctimes_ik(b, -P.f * (double)n);
      }
    });
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = mk(-e1[m] * u[m], -e2[m] * u[m]);
  {
    cd* o10 = A.O[10] + fo;
    cd* o12 = A.O[12] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
      if (n < Jt) { o10[TF * n] = a; o12[TF * n] = b; }
    });
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = mk(-e1[m] * w[m], -e2[m] * w[m]);
  {
    cd* o11 = A.O[11] + fo;
    cd* o13 = A.O[13] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
      if (n < Jt) {
        o11[TF * n] = ctimes_ik(a, -P.f * (double)n);
        o13[TF * n] = ctimes_ik(b, -P.f * (double)n);
      }
    });
  }
  // ubar = -xi w_x - eta_1 c1x - eta_2 c2x, vbar = -xi w_y - eta_1 c1y - eta_2 c2y
  double ub[8], vb[8];
  ld.template next<false, true, true, false>(A.Hwx, A.Hwy, A.Hdx1, A.Hc1, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    ub[m] = -xi[m] * (v[m].x * sc);
    vb[m] = -xi[m] * (v[m].y * sc);
  }
  ld.template next<true, true, true, false>(A.Hdx1, A.Hc1, A.Hdx2, A.Hc2, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    ub[m] += -e1[m] * (v[m].x * sc);
    vb[m] += -e1[m] * (v[m].y * sc);
  }
  ld.template next<true, true, false, true>(A.Hdx2, A.Hc2, nullptr, nullptr, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    ub[m] += -e2[m] * (v[m].x * sc);
    vb[m] += -e2[m] * (v[m].y * sc);
    v[m] = mk(ub[m], vb[m]);
  }
  {
    cd* o0 = A.O[0] + fo;
    cd* o1 = A.O[1] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
      if (n < Jt) {
        o0[TF * n] = ctimes_ik(a, -P.f * (double)n);   // -i ky R(ubar)
        o1[TF * n] = b;                                 // R(vbar)
      }
    });
  }
}

// ============================================================ adj_rowD2
// Transposes of the surface-tension force F = A1 grad c1 + A2 grad c2 with
// Fxbar = d_y xi, Fybar = -d_x xi: Abar_i = d_y xi c_ix - d_x xi c_iy,
// mubar (A_i = A_i1 mu_1 + A_i2 mu_2), and the grad-c bars d_y xi A_i, -d_x xi A_i.
// Outputs: O4 = R(c1x bar), O5 = -i ky R(c1y bar) + mc1 ky^2 R(mu1bar),
// O6 = R(mu1bar), O7, O8, O9 likewise for c2; partials A11..A22 = sum Abar_i mu_j.
// grid: N / R
template <int LOG2N>
__global__ __launch_bounds__(RowCfg<LOG2N>::NT, 2) void k_adj_rowD2(AdjArgs args) {
  const AdjBufs& A = args.A;
  const KParams& P = args.P;
  using RC = RowCfg<LOG2N>;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  constexpr bool STG = AdjRowLoader<LOG2N>::STG;
  __shared__ cd stg_all[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;
  const int x = row_of<LOG2N>(r);
  cd* lds = lds_all + r;
  const size_t fo = tix<TF>(x, 0, P.ldJ);
  const double sc = P.scale;
  const int Jt = P.Jt;
  const AdjRowLoader<LOG2N> ld{{stg_all, Jt}, tix<TJ>(x, 0, P.ldJ), Jt, P.f};
  cd v[8];
  // twiddles loaded per transform: held ones (as in adj_rowD1) spill 48
  // VGPRs here, against 16, and measured slower
  double xix[8], xiy[8], Ab1[8], Ab2[8];
  ld.template first<true, true, true, true>(A.Hxix, A.Hxi, A.Hdx1, A.Hc1, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, A.tw, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) { xix[m] = v[m].x * sc; xiy[m] = v[m].y * sc; }
  ld.template next<true, true, true, false>(A.Hdx1, A.Hc1, A.Hdx2, A.Hc2, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, A.tw, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) Ab1[m] = xiy[m] * (v[m].x * sc) - xix[m] * (v[m].y * sc);
  ld.template next<true, true, true, false>(A.Hdx2, A.Hc2, A.Hmu1, A.Hmu2, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, A.tw, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) Ab2[m] = xiy[m] * (v[m].x * sc) - xix[m] * (v[m].y * sc);
  ld.template next<false, true, false, true>(A.Hmu1, A.Hmu2, nullptr, nullptr, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, A.tw, t);
  double sq[4] = {0.0, 0.0, 0.0, 0.0};
  double A1[8], A2[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const double m1 = v[m].x * sc, m2 = v[m].y * sc;
    A1[m] = (P.A11 * m1 + P.A12 * m2);   // surface_tension.c:144
    A2[m] = (P.A21 * m1 + P.A22 * m2);   // surface_tension.c:195
    sq[0] += Ab1[m] * m1;
    sq[1] += Ab1[m] * m2;
    sq[2] += Ab2[m] * m1;
    sq[3] += Ab2[m] * m2;
    v[m] = mk(xiy[m] * A1[m], xiy[m] * A2[m]);
  }
  {
    cd* o4 = A.O[4] + fo;
    cd* o7 = A.O[7] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, A.tw, t, [&](int n, cd a, cd b) {
      if (n < Jt) { o4[TF * n] = a; o7[TF * n] = b; }
    });
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = mk(-xix[m] * A1[m], P.A11 * Ab1[m] + P.A21 * Ab2[m]);
  {
    cd* o5 = A.O[5] + fo;
    cd* o6 = A.O[6] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, A.tw, t, [&](int n, cd a, cd b) {
      if (n < Jt) {
        const double ky = P.f * (double)n;
        o5[TF * n] = cadd(ctimes_ik(a, -ky), cscale(b, P.mc1 * __dmul_rn(ky, ky)));
        o6[TF * n] = b;
      }
    });
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = mk(-xix[m] * A2[m], P.A12 * Ab1[m] + P.A22 * Ab2[m]);
  {
    cd* o8 = A.O[8] + fo;
    cd* o9 = A.O[9] + fo;
    rfft_split<LOG2N, RC::R>(v, lds, A.tw, t, [&](int n, cd a, cd b) {
      if (n < Jt) {
        const double ky = P.f * (double)n;
        o8[TF * n] = cadd(ctimes_ik(a, -ky), cscale(b, P.mc2 * __dmul_rn(ky, ky)));
        o9[TF * n] = b;
      }
    });
  }
  block_sum<4, RC::NT>(sq, (double*)lds_all, A.part + (size_t)AQ_A11 * A.mb + blockIdx.x, A.mb);
}

// ============================================================ adj_colBC
// grid: 1-D over (tile, i) items, i = phase field; tiles cover the Jt kept columns.
template <int LOG2N>
__global__ __launch_bounds__(ColCfg<LOG2N>::NT) void k_adj_colBC(AdjArgs args) {
  const AdjBufs& A = args.A;
  const KParams& P = args.P;
  using CF = ColCfg<LOG2N>;
  constexpr int T = CF::T;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  int tile, g;
  double sq[4] = {0.0, 0.0, 0.0, 0.0};   // mc1, mc2, lg1, lg2
  double* dst = A.part + blockIdx.x;
  if (!col_item((P.Jt + CF::C - 1) / CF::C, 2, &tile, &g)) {
    // padding workgroup: its partial slots still hold zeros
    block_sum<2, CF::NT>(sq, (double*)lds_all, dst + (size_t)AQ_MC1 * A.mb, A.mb);
    block_sum<2, CF::NT>(sq + 2, (double*)lds_all, dst + (size_t)AQ_LG1 * A.mb, A.mb);
    return;
  }
  const int j = tile * CF::C + c;
  const bool ok = j < P.Jt;
  const int a = ok ? A.amask[j] : -1;
  cd* lds = lds_all + c;
  COL_TWIDDLES(A.tw);
  const double ky = P.f * (double)j;
  const double wj = (j == 0) ? 1.0 : 2.0;   // half-spectrum weight (j = N/2 is never kept)
  cd v[CF::P], s[CF::P], gi[CF::P];
  // Mbar_i = T(R(mubar_i)) - (lambda/g_i) k^2 g_i
  col_load_rm<LOG2N, CF::P, TF>(A.O[g == 0 ? 6 : 9], P.ldJ, j, ok, t, v);
  fft<LOG2N, CF::P, false, CF::C, false, LT>(v, lds, twp, t);
  col_load_masked<LOG2N, CF::P>(g == 0 ? A.C1 : A.C2, j, a, t, s);
  col_load_masked<LOG2N, CF::P>(A.Gi[1 + g], j, a, t, gi);
  const double lg = g == 0 ? P.lg1 : P.lg2;
  double smc = 0.0;
#pragma unroll
  for (int m = 0; m < CF::P; ++m) {
    const int i = t + m * T;
    const double ks = ksq(kxof<LOG2N>(i, P.f), ky);
    if (absidx<LOG2N>(i) <= a) {
      smc += wj * ks * (v[m].x * s[m].x + v[m].y * s[m].y);
      v[m] = csub(v[m], cscale(gi[m], lg * ks));
    } else {
      v[m] = mk(0.0, 0.0);
    }
  }
  // lg partial: -<g_i, k^2 M_i> with M_i = the column FFT of pass B's output
  col_load_rm<LOG2N, CF::P, TF>(g == 0 ? A.Mh1 : A.Mh2, P.ldF, j, ok, t, s);
  fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
  double slg = 0.0;
#pragma unroll
  for (int m = 0; m < CF::P; ++m) {
    const double ks = ksq(kxof<LOG2N>(t + m * T, P.f), ky);
    slg -= wj * ks * (gi[m].x * s[m].x + gi[m].y * s[m].y);   // gi is zero off the kept modes
  }
  fft<LOG2N, CF::P, true, CF::C, true, LT>(v, lds, twp, t);
  col_store_rm<LOG2N, CF::P, TJ>(g == 0 ? A.Hmb1 : A.Hmb2, P.ldJ, j, ok, t, v);
  sq[g] = smc * P.scale;
  sq[2 + g] = slg * P.scale;
  block_sum<2, CF::NT>(sq, (double*)lds_all, dst + (size_t)AQ_MC1 * A.mb, A.mb);
  block_sum<2, CF::NT>(sq + 2, (double*)lds_all, dst + (size_t)AQ_LG1 * A.mb, A.mb);
}

// ============================================================ adj_rowB
// grid: N / R
template <int LOG2N>
__global__ __launch_bounds__(RowCfg<LOG2N>::NT) void k_adj_rowB(AdjArgs args) {
  const AdjBufs& A = args.A;
  const KParams& P = args.P;
  using RC = RowCfg<LOG2N>;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;
  const int x = row_of<LOG2N>(r);
  cd* lds = lds_all + r;
  const size_t ro = tix<TJ>(x, 0, P.ldJ);
  const double sc = P.scale;
  cd v[8];
  RowTw<LOG2N, RC::R> tws;
  tws.init(A.tw, t);
  double c1[8], c2[8];
  row_load_pair<LOG2N, false>(A.Hc1 + ro, A.Hc2 + ro, P.Jt, P.f, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) { c1[m] = v[m].x * sc; c2[m] = v[m].y * sc; }
  row_load_pair<LOG2N, false>(A.Hmb1 + ro, A.Hmb2 + ro, P.Jt, P.f, t, v);
  rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
  // fac, G1, G2, G3, g1i, g2i, g3i
  double sq[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const double r1 = v[m].x * sc, r2 = v[m].y * sc;   // muF bars
    const double a1 = c1[m], a2 = c2[m];
    const double a3 = (1.0 - a1 - a2);
    const double c123 = a1 * a2 * a3;
    const double fc1 = a1 * (1.0 - a1) * (1.0 - 2.0 * a1);
    const double fc2 = a2 * (1.0 - a2) * (1.0 - 2.0 * a2);
    const double fc3 = a3 * (1.0 - a3) * (1.0 - 2.0 * a3);
    const double G1 = P.g1 * fc1 + P.X6 * c123 * a2 * a3;
    const double G2 = P.g2 * fc2 + P.X6 * c123 * a1 * a3;
    const double G3 = P.g3 * fc3 + P.X6 * c123 * a1 * a2;
    const double m1 = (G1 - G2) * P.g2i + (G1 - G3) * P.g3i;
    const double m2 = (G2 - G1) * P.g1i + (G2 - G3) * P.g3i;
    sq[0] += r1 * m1 + r2 * m2;
    sq[4] += P.fac * r2 * (G2 - G1);
    sq[5] += P.fac * r1 * (G1 - G2);
    sq[6] += P.fac * (r1 * (G1 - G3) + r2 * (G2 - G3));
    const double Gb1 = P.fac * (r1 * (P.g2i + P.g3i) - r2 * P.g1i);
    const double Gb2 = P.fac * (-r1 * P.g2i + r2 * (P.g1i + P.g3i));
    const double Gb3 = -P.fac * (r1 + r2) * P.g3i;
    sq[1] += Gb1 * fc1;
    sq[2] += Gb2 * fc2;
    sq[3] += Gb3 * fc3;
    const double X2 = 2.0 * P.X6;
    const double fp1 = 1.0 - 6.0 * a1 + 6.0 * a1 * a1;
    const double fp2 = 1.0 - 6.0 * a2 + 6.0 * a2 * a2;
    const double fp3 = 1.0 - 6.0 * a3 + 6.0 * a3 * a3;
    // dG_j/dc_k with c1, c2, c3 independent; then c3 = 1 - c1 - c2
    const double d1 = Gb1 * (P.g1 * fp1 + P.X6 * a2 * a2 * a3 * a3) + Gb2 * (X2 * a1 * a2 * a3 * a3) +
                      Gb3 * (X2 * a1 * a2 * a2 * a3);
    const double d2 = Gb1 * (X2 * a1 * a2 * a3 * a3) + Gb2 * (P.g2 * fp2 + P.X6 * a1 * a1 * a3 * a3) +
                      Gb3 * (X2 * a1 * a1 * a2 * a3);
    const double d3 = Gb1 * (X2 * a1 * a2 * a2 * a3) + Gb2 * (X2 * a1 * a1 * a2 * a3) +
                      Gb3 * (P.g3 * fp3 + P.X6 * a1 * a1 * a2 * a2);
    v[m] = mk(d1 - d3, d2 - d3);
  }
    cd* o1 = A.Ocb1 + tix<TF>(x, 0, P.ldJ);
  cd* o2 = A.Ocb2 + tix<TF>(x, 0, P.ldJ);
  const int Jt = P.Jt;
  rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
    if (n < Jt) { o1[TF * n] = a; o2[TF * n] = b; }
  });
  block_sum<7, RC::NT>(sq, (double*)lds_all, A.part + blockIdx.x, A.mb);
}

// ============================================================ adj_colE
// STAGE 2: product at a (abar = ubar + J, next input Z1 abar - Z2 ubar);
// STAGE 1: product at u (ubar = E abar + J, next input Z2 ubar);
// STAGE 0: no product (next input Z2 ubar), used once after the seed.
// Every stage ends with the column iFFTs of the next input for adj_rowD.
// grid: 1-D over (tile, field) items: omega, c1, c2; tiles cover the Jt kept columns.
template <int LOG2N, int STAGE>
__global__ __launch_bounds__(ColCfg<LOG2N>::NT) void k_adj_colE(AdjArgs args) {
  const AdjBufs& A = args.A;
  const KParams& P = args.P;
  using CF = ColCfg<LOG2N>;
  constexpr int N = CF::N, T = CF::T, H = N / 2 + 1;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  int tile, g;
  if (!col_item((P.Jt + CF::C - 1) / CF::C, 3, &tile, &g)) return;
  const int j = tile * CF::C + c;
  const bool ok = j < P.Jt;
  const int a = ok ? A.amask[j] : -1;
  cd* lds = lds_all + c;
  COL_TWIDDLES(A.tw);
  const double ky = P.f * (double)j;
  cd v[CF::P], s[CF::P];
  if constexpr (STAGE != 0) {
    if (g == 0) {
      // omega: T[ Q (R(-i ky ubar) + i kx R(vbar)) - i kx R(w_x bar) - i ky R(w_y bar) ]
      col_load_rm<LOG2N, CF::P, TF>(A.O[0], P.ldJ, j, ok, t, v);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(v, lds, twp, t);
      col_load_rm<LOG2N, CF::P, TF>(A.O[1], P.ldJ, j, ok, t, s);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) {
        const double kx = kxof<LOG2N>(t + m * T, P.f);
        const double ks = ksq(kx, ky);
        const cd q = cadd(v[m], ctimes_ik(s[m], kx));
        v[m] = ks > 1e-12 ? mk(q.x / ks, q.y / ks) : mk(0.0, 0.0);   // psi = w / k^2 (velocity.c:20-25)
      }
      col_load_rm<LOG2N, CF::P, TF>(A.O[2], P.ldJ, j, ok, t, s);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = cadd(v[m], ctimes_ik(s[m], -kxof<LOG2N>(t + m * T, P.f)));
      col_load_rm<LOG2N, CF::P, TF>(A.O[3], P.ldJ, j, ok, t, s);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = cadd(v[m], s[m]);
    } else {
      // c_i: T[ R(cbar_i) - i kx R(c_ix bar) - i ky R(c_iy bar) + mc_i k^2 R(mubar_i) ]
      const int o = g == 1 ? 4 : 7, oa = g == 1 ? 10 : 12;
      col_load_rm<LOG2N, CF::P, TF>(A.O[o], P.ldJ, j, ok, t, s);
      {
        cd w[CF::P];
        col_load_rm<LOG2N, CF::P, TF>(A.O[oa], P.ldJ, j, ok, t, w);
#pragma unroll
        for (int m = 0; m < CF::P; ++m) s[m] = cadd(s[m], w[m]);
      }
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], -kxof<LOG2N>(t + m * T, P.f));
      col_load_rm<LOG2N, CF::P, TF>(A.O[o + 1], P.ldJ, j, ok, t, s);
      {
        cd w[CF::P];
        col_load_rm<LOG2N, CF::P, TF>(g == 1 ? A.Ocb1 : A.Ocb2, P.ldJ, j, ok, t, w);
#pragma unroll
        for (int m = 0; m < CF::P; ++m) s[m] = cadd(s[m], w[m]);
        col_load_rm<LOG2N, CF::P, TF>(A.O[oa + 1], P.ldJ, j, ok, t, w);
#pragma unroll
        for (int m = 0; m < CF::P; ++m) s[m] = cadd(s[m], w[m]);
      }
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = cadd(v[m], s[m]);
      col_load_rm<LOG2N, CF::P, TF>(A.O[o + 2], P.ldJ, j, ok, t, s);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
      const double mc = g == 1 ? P.mc1 : P.mc2;
#pragma unroll
      for (int m = 0; m < CF::P; ++m) {
        const double kx = kxof<LOG2N>(t + m * T, P.f);
        v[m] = cadd(v[m], cscale(s[m], mc * __dmul_rn(kx, kx)));
      }
    }
  }
  // ETD2RK transpose on the kept modes; s[] receives the next product's input
  cd* Ub = A.Ub[g];
  cd* Ab = A.Ab[g];
  const cd* tab1 = g == 0 ? A.tabw1 : A.tabc1;
  const double* tab2 = g == 0 ? A.tabw2 : A.tabc2;
#pragma unroll
  for (int m = 0; m < CF::P; ++m) {
    const int i = t + m * T;
    const bool kept = absidx<LOG2N>(i) <= a;
    cd gn = mk(0.0, 0.0);
    if (kept) {
      const size_t idx = (size_t)j * N + i;
      const int ti = j * H + absidx<LOG2N>(i);
      if constexpr (STAGE == 2) {
        const cd ub = Ub[idx];
        const double z2 = tab2[ti];
        const cd ab = cadd(ub, v[m]);
        Ab[idx] = ab;
        gn = csub(cscale(ab, tab1[ti].y), cscale(ub, z2));
      } else if constexpr (STAGE == 1) {
        const cd ab = Ab[idx];
        const cd ub = cadd(cscale(ab, tab1[ti].x), v[m]);
        Ub[idx] = ub;
        gn = cscale(ub, tab2[ti]);
      } else {
        gn = cscale(Ub[idx], tab2[ti]);
      }
      if (g > 0) A.Gi[g][idx] = gn;
    }
    s[m] = gn;
  }
  // column iFFTs of the next input: xi, d_x xi (omega) or eta_i
  if (g == 0) {
#pragma unroll
    for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], kxof<LOG2N>(t + m * T, P.f));
    fft<LOG2N, CF::P, true, CF::C, true, LT>(s, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(A.Hxi, P.ldJ, j, ok, t, s);
    fft<LOG2N, CF::P, true, CF::C, true, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(A.Hxix, P.ldJ, j, ok, t, v);
  } else {
    fft<LOG2N, CF::P, true, CF::C, true, LT>(s, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(g == 1 ? A.Het1 : A.Het2, P.ldJ, j, ok, t, s);
  }
}

// Spectral adjoint of the c1 shell-spectrum loss at the current state:
// c1bar = 4 (S(s(k)) - S_target(s(k))) c1(k) / N^2 on the kept modes, zero
// elsewhere; omega bar = c2 bar = 0.  dS[s] = S(s) - S_target(s), nsh shells.
// grid: ceil(Ny2*N / 256)
static __global__ void k_adj_seed_spectrum(const cd* __restrict__ C1, cd* Ub0, cd* Ub1, cd* Ub2,
                                           const double* __restrict__ dS, const int* __restrict__ amask,
                                           int N, double f, double dk, double coef, int nsh) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int H = N / 2 + 1;
  if (e >= (size_t)N * H) return;
  const int j = (int)(e / N), i = (int)(e % N);
  const int ai = i <= N / 2 ? i : N - i;
  cd r = mk(0.0, 0.0);
  if (ai <= amask[j]) {
    const double kx = f * (double)(i <= N / 2 ? i : i - N);
    const double ky = f * (double)j;
    const double ks = __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
    int ik = (int)((sqrt(ks) / dk) + 0.5);   // spectrum.c shell index
    ik = ik < nsh ? ik : nsh - 1;
    r = cscale(C1[e], coef * dS[ik]);
  }
  Ub0[e] = mk(0.0, 0.0);
  Ub1[e] = r;
  Ub2[e] = mk(0.0, 0.0);
}

// Real-space adjoint of the ternary free energy ephi + egphi
// (free_energy.c:238-248,317-335) of the fields f = c1, c2, d/dx c1, d/dy c1,
// d/dx c2, d/dy c2 (packed [N][N] each), overwritten in place by
// dL/dc1, dL/dc2, dL/d(d/dx c1), ...: with s = c1 + c2 = 1 - c3,
// ephi = b sum g1 c1^2 (1-c1)^2 + g2 c2^2 (1-c2)^2 + g3 s^2 (1-s)^2 and
// egphi = q sum g1 |grad c1|^2 + g2 |grad c2|^2 + g3 |grad c1 + grad c2|^2,
// b = 6 scale/eps, q = 0.375 eps scale.  grid: ceil(N*N / 256)
static __global__ void k_fe_adjoint(double* c1, double* c2, double* c1x, double* c1y, double* c2x, double* c2y,
                                    size_t n, double g1, double g2, double g3, double b, double q) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double a1 = c1[p], a2 = c2[p], sm = a1 + a2;
  const double d3 = g3 * 2.0 * sm * (1.0 - sm) * (1.0 - 2.0 * sm);   // d/ds of g3 s^2 (1-s)^2
  c1[p] = b * (g1 * 2.0 * a1 * (1.0 - a1) * (1.0 - 2.0 * a1) + d3);
  c2[p] = b * (g2 * 2.0 * a2 * (1.0 - a2) * (1.0 - 2.0 * a2) + d3);
  const double sx = c1x[p] + c2x[p], sy = c1y[p] + c2y[p];
  const double u1x = c1x[p], u1y = c1y[p], u2x = c2x[p], u2y = c2y[p];
  c1x[p] = q * 2.0 * (g1 * u1x + g3 * sx);
  c1y[p] = q * 2.0 * (g1 * u1y + g3 * sy);
  c2x[p] = q * 2.0 * (g2 * u2x + g3 * sx);
  c2y[p] = q * 2.0 * (g2 * u2y + g3 * sy);
}

// Spectral seed from the r2c of those adjoints (X[k], column-major, already
// truncated): c_i bar = X_c - i kx X_x - i ky X_y (the transpose of the
// multipliers i kx, i ky is their conjugate); omega bar = 0.
// grid: ceil(Ny2*N / 256)
static __global__ void k_fe_seed(const cd* __restrict__ X1, const cd* __restrict__ X2,
                                 const cd* __restrict__ X1x, const cd* __restrict__ X1y,
                                 const cd* __restrict__ X2x, const cd* __restrict__ X2y, cd* Ub0, cd* Ub1, cd* Ub2,
                                 int N, double f) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int H = N / 2 + 1;
  if (e >= (size_t)N * H) return;
  const int j = (int)(e / N), i = (int)(e % N);
  const double kx = f * (double)(i <= N / 2 ? i : i - N), ky = f * (double)j;
  // (-i k) z = (k Im z, -k Re z)
  const cd a = X1x[e], bb = X1y[e], c = X2x[e], d = X2y[e];
  Ub0[e] = mk(0.0, 0.0);
  Ub1[e] = mk(X1[e].x + kx * a.y + ky * bb.y, X1[e].y - kx * a.x - ky * bb.x);
  Ub2[e] = mk(X2[e].x + kx * c.y + ky * d.y, X2[e].y - kx * c.x - ky * d.x);
}

// acc[q] += sum over slots and workgroups of part[slot][q][b] (b < nb[q]),
// in a fixed order.  grid: ADJ_NQ workgroups of 256 threads.
static __global__ void k_adj_accum(const double* __restrict__ part, int nslot, size_t slot_stride, int mb,
                                   const int* __restrict__ nbq, double* acc) {
  __shared__ double red[256];
  const int q = blockIdx.x;
  const int nb = nbq[q];
  double s = 0.0;
  for (int sl = 0; sl < nslot; ++sl)
    for (int b = threadIdx.x; b < nb; b += 256) s += part[sl * slot_stride + (size_t)q * mb + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[q] += red[0];
}

}  // namespace chns

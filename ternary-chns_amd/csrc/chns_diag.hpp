// Device diagnostics of the spectral state (free_energy.c, kinetic_energy.c,
// perimeter.c, spectrum.c) and the real-field dump path (main.c:332-358).
//
// Real-space fields come from a full (not dealiased unless asked) 2-D c2r:
// column iFFT of every ky column (k_col_inv_full), then a paired row c2r
// (k_row_c2r_full) into packed [N][N] real buffers.  Reductions run in a
// fixed order (per-block trees, then one block over the block partials), so
// results are bitwise repeatable.
#pragma once
#include "chns_kernels.hpp"

namespace chns {


// Column iFFT of a whole state field with a per-mode operator, optionally
// truncated; writes the TJ-row tiled [N][ldF] array.  grid: 1-D items.
template <int LOG2N, int OP, bool TRUNC>
__global__ __launch_bounds__(ColCfg<LOG2N>::NT) void k_col_inv_full(const cd* __restrict__ S, cd* out,
                                                                  const cd* __restrict__ tw,
                                                                  const int* __restrict__ amask,
                                                                  const double* __restrict__ ft, int ldF,
                                                                  double f) {
  using CF = ColCfg<LOG2N>;
  constexpr int N = CF::N, T = CF::T, H = N / 2 + 1;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  int tile, g;
  if (!col_item((H + CF::C - 1) / CF::C, 1, &tile, &g)) return;
  const int j = tile * CF::C + c;
  const bool ok = j < H;
  cd* lds = lds_all + c;
  const int a = ok ? (TRUNC ? amask[j] : N) : -1;
  const double ky = f * (double)j;
  cd v[CF::P];
#pragma unroll
  for (int m = 0; m < CF::P; ++m) {
    const int i = t + m * T;
    cd z = (absidx<LOG2N>(i) <= a) ? S[(size_t)j * N + i] : mk(0.0, 0.0);
    const double kx = kxof<LOG2N>(i, f);
    if (OP == COP_IKX) z = ctimes_ik(z, kx);                       // I*kx*c
    if (OP == COP_IKY) z = ctimes_ik(z, ky);                       // I*ky*c
    if (OP == COP_U || OP == COP_V) {                              // velocity.c:20-39
      const double ks = ksq(kx, ky);
      z = ks > 1e-12 ? mk(z.x / ks, z.y / ks) : mk(0.0, 0.0);
      z = OP == COP_U ? ctimes_ik(z, ky) : ctimes_ik(z, -kx);
    }
    if (TRUNC && ft && ok) z = filt2(z, filt_at<LOG2N>(ft, j, i));   // SMOOTH_FILTER truncation
    v[m] = z;
  }
  fft<LOG2N, CF::P, true, CF::C>(v, lds, tw, t);
  col_store_rm<LOG2N, CF::P, TJ>(out, ldF, j, ok, t, v);
}

// Paired row c2r of full half-spectra (all ky columns; real parts at ky = 0
// and ky = N/2 as a c2r uses them), scaled by 1/(Nx*Ny) (fft.c:21-29),
// into packed real [N][N] rows.  grid: N / R
template <int LOG2N>
__global__ __launch_bounds__(RowCfg<LOG2N>::NT) void k_row_c2r_full(const cd* __restrict__ X,
                                                                  const cd* __restrict__ Y, double* oa,
                                                                  double* ob, const cd* __restrict__ tw,
                                                                  int ldF, double scale) {
  using RC = RowCfg<LOG2N>;
  constexpr int N = RC::N, T = RC::T, H = N / 2 + 1;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;
  const int x = row_of<LOG2N>(r);
  cd* lds = lds_all + r;
  cd v[8];
  const size_t ro = tix<TJ>(x, 0, ldF);
  row_load_pair<LOG2N, false>(X + ro, Y ? Y + ro : nullptr, H, 0.0, t, v);
  rfft<LOG2N, true, RC::R>(v, lds, tw, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const size_t o = (size_t)x * N + rpos<LOG2N, RC::R>(t, m);
    oa[o] = v[m].x * scale;
    if (ob) ob[o] = v[m].y * scale;
  }
}

// ---------------------------------------------------------------- reductions
// Block sums of up to 4 quantities over packed real fields; each block covers
// RB = 4096 points (256 threads x 16), pairwise tree in LDS.  The partials
// [nblocks][4] are summed by k_reduce_final in block order.
constexpr int RB_THREADS = 256, RB_PTS = 16;



template <int KIND>
__device__ __forceinline__ void red_point(const RedArgs& A, size_t p, double* q) {
  if constexpr (KIND == RED_FE_TER) {
    // free_energy.c:238-248, 317-331: bulk and gradient energies (scaled on the host)
    const double c1 = A.f[0][p], c2 = A.f[1][p], c3 = 1.0 - c1 - c2;
    q[0] += (A.g1 * c1 * c1 * (1.0 - c1) * (1.0 - c1) + A.g2 * c2 * c2 * (1.0 - c2) * (1.0 - c2) +
             A.g3 * c3 * c3 * (c1 + c2) * (c1 + c2));
    const double c1x = A.f[2][p], c1y = A.f[3][p], c2x = A.f[4][p], c2y = A.f[5][p];
    q[1] += (A.g1 * (c1x * c1x + c1y * c1y) + A.g2 * (c2x * c2x + c2y * c2y) + A.g3 * (c1x + c2x) * (c1x + c2x) +
             A.g3 * (c1y + c2y) * (c1y + c2y));
  } else if constexpr (KIND == RED_FE_BIN) {
    // free_energy.c:231-236, 283-285
    const double c = A.f[0][p];
    q[0] += c * c * (1.0 - c) * (1.0 - c);
    const double cx = A.f[1][p], cy = A.f[2][p];
    q[1] += (cx * cx + cy * cy);
  } else if constexpr (KIND == RED_KE_TER || KIND == RED_KE_BIN || KIND == RED_KE_NS) {
    // kinetic_energy.c:46-90
    const double cx = A.f[0][p], cy = A.f[1][p];
    const double e = 0.5 * (cx * cx + cy * cy);
    q[0] += e;
    if constexpr (KIND != RED_KE_NS) {
      if (A.f[2][p] >= 0.5) { q[1] += e; q[2] += 1.0; }
    }
    if constexpr (KIND == RED_KE_TER) {
      if (A.f[3][p] >= 0.5) { q[3] += e; q[4] += 1.0; }
    }
  } else {
    // perimeter.c:136-238: sign change to the x+1 or y+1 neighbour (periodic)
    const int N = A.N;
    const int x = (int)(p / N), y = (int)(p % N);
    const size_t px = (size_t)((x + 1) & (N - 1)) * N + y, py = (size_t)x * N + ((y + 1) & (N - 1));
#pragma unroll
    for (int fi = 0; fi < 2; ++fi) {
      const double* c = A.f[fi];
      if (!c) continue;
      const double s0 = c[p] >= 0.5 ? 1.0 : -1.0;
      const double sx = c[px] >= 0.5 ? 1.0 : -1.0;
      const double sy = c[py] >= 0.5 ? 1.0 : -1.0;
      if ((s0 * sx) < 0.0 || (s0 * sy) < 0.0) q[2 * fi] += 1.0;
      if (s0 > 0.0) q[2 * fi + 1] += 1.0;
    }
  }
}



template <int KIND>
__global__ __launch_bounds__(RB_THREADS) void k_reduce(RedArgs A, double* partial) {
  __shared__ double sh[RED_Q][RB_THREADS];
  double q[RED_Q] = {0, 0, 0, 0, 0};
  const size_t base = (size_t)blockIdx.x * RB_THREADS * RB_PTS;
#pragma unroll 4
  for (int k = 0; k < RB_PTS; ++k) {
    const size_t p = base + (size_t)k * RB_THREADS + threadIdx.x;
    if (p < (size_t)A.N * A.N) red_point<KIND>(A, p, q);
  }
#pragma unroll
  for (int i = 0; i < RED_Q; ++i) sh[i][threadIdx.x] = q[i];
  __syncthreads();
  for (int w = RB_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int i = 0; i < RED_Q; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < RED_Q) partial[(size_t)blockIdx.x * RED_Q + threadIdx.x] = sh[threadIdx.x][0];
}

// Sum nb partial rows in a fixed pairwise order; one block of 256 threads.
__global__ __launch_bounds__(RB_THREADS) void k_reduce_final(const double* __restrict__ partial, int nb,
                                                             double* out) {
  __shared__ double sh[RED_Q][RB_THREADS];
  double q[RED_Q] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += RB_THREADS)
#pragma unroll
    for (int i = 0; i < RED_Q; ++i) q[i] += partial[(size_t)b * RED_Q + i];
#pragma unroll
  for (int i = 0; i < RED_Q; ++i) sh[i][threadIdx.x] = q[i];
  __syncthreads();
  for (int w = RB_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int i = 0; i < RED_Q; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < RED_Q) out[threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------- spectrum
// spectrum.c:52-121.  One thread per kx row i walks ky = 0..N/2; shells
// ik = (int)(|k|/dk + 0.5) are non-decreasing along the row and step by at
// most one, so row i covers every shell of [ik(i,0), ik(i,N/2)] once.  It
// writes each shell's run sums into shell-major partials part[q][s][i]
// (quantities 0 ek, 1 ensk, 2 sph1 (sph), 3 sph2, 4 sp12) and party[b][i]
// (eky, b = iky); entries outside a row's range are never written, and the
// reducer skips them, so nothing needs clearing.  ekx per row is one number.
__device__ __forceinline__ int spec_shell(double kxsqr, double ky, double dk) {
  const double ks = __dadd_rn(kxsqr, __dmul_rn(ky, ky));
  return (int)((sqrt(ks) / dk) + 0.5);
}

__global__ void k_spec_rows(SpecArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N) return;
  const int N = A.N, H = N / 2 + 1;
  const int ai = i <= N / 2 ? i : N - i;
  const double kx = A.f * (double)(i <= N / 2 ? i : i - N);
  const double kxsqr = __dmul_rn(kx, kx);
  double acc[5] = {0, 0, 0, 0, 0};
  int cur = -1;
  double ex = 0.0;
  auto flush = [&](int s) {
    if (s >= 0 && s < A.nsh)
      for (int q = 0; q < 5; ++q) A.part[((size_t)q * A.nsh + s) * N + i] = acc[q];
    for (int q = 0; q < 5; ++q) acc[q] = 0.0;
  };
  for (int j = 0; j < H; ++j) {
    const double ky = A.f * (double)j;
    const double kysqr = __dmul_rn(ky, ky);
    const double ks = __dadd_rn(kxsqr, kysqr);
    const int ik = (int)((sqrt(ks) / A.dk) + 0.5);
    const int iky = (int)((sqrt(kysqr) / A.dk) + 0.5);
    if (ik != cur) {
      flush(cur);
      // shells step by at most one along a row (|k| grows by at most dk per
      // column); a skipped shell would still get its zero entries
      for (int sk = cur + 1; cur >= 0 && sk < ik; ++sk) flush(sk);
      cur = ik;
    }
    const bool kept = ai <= A.amask[j];   // truncation of u and the phase fields (spectrum.c:22-27)
    const double fe = A.filt ? A.filt[(size_t)j * H + ai] : 1.0;   // ... or the smooth filter
    const size_t idx = (size_t)j * N + i;
    double eu = 0.0;
    if (kept) {
      // get_velocity (velocity.c:20-39)
      cd w = A.W[idx];
      w = ks > 1e-12 ? mk(w.x / ks, w.y / ks) : mk(0.0, 0.0);
      cd ux = ctimes_ik(w, ky), uy = ctimes_ik(w, -kx);
      if (A.filt) {
        ux = filt2(ux, fe);
        uy = filt2(uy, fe);
      }
      eu = (ux.x * ux.x + ux.y * ux.y + uy.x * uy.x + uy.y * uy.y) * A.scale2;
    }
    const double ens = eu * ks;
    const bool edge = (ky == 0.0 || ky == A.kymax);
    const double wh = edge ? 0.5 : 1.0, w2 = edge ? 1.0 : 2.0;
    acc[0] += wh * eu;
    acc[1] += wh * ens;
    ex += wh * eu;
    if (iky < A.nxy) A.party[(size_t)iky * N + i] = wh * eu;
    if (A.nph >= 1 && kept) {
      const cd c1 = A.filt ? filt2(A.C1[idx], fe) : A.C1[idx];
      acc[2] += w2 * ((c1.x * c1.x + c1.y * c1.y) * A.scale2);
      if (A.nph >= 2) {
        const cd c2 = A.filt ? filt2(A.C2[idx], fe) : A.C2[idx];
        acc[3] += w2 * ((c2.x * c2.x + c2.y * c2.y) * A.scale2);
        acc[4] += w2 * fabs((c1.x * c2.x + c1.y * c2.y) * A.scale2);
      }
    }
  }
  flush(cur);
  A.partx[i] = ex;
}

// Workgroup b < nsh: out[q*nsh + b] = sum over rows i (ascending, in a fixed
// tree) of part[q][b][i] for the rows whose shell range holds b.  Workgroup
// nsh + b: outy[b] = sum_i party[b][i] (the rows write every b <= N/2) and
// outx[b] = sum over rows with ikx(i) == b of partx[i].  256 threads.
__global__ void k_spec_sum(SpecArgs A, double* out, double* outy, double* outx) {
  __shared__ double red[5][256];
  const int N = A.N, H = N / 2 + 1, tid = threadIdx.x;
  double acc[5] = {0, 0, 0, 0, 0};
  const bool shell = (int)blockIdx.x < A.nsh;
  const int b = shell ? (int)blockIdx.x : (int)blockIdx.x - A.nsh;
  for (int i = tid; i < N; i += 256) {
    const double kx = A.f * (double)(i <= N / 2 ? i : i - N);
    const double kxsqr = __dmul_rn(kx, kx);
    if (shell) {
      const int lo = spec_shell(kxsqr, 0.0, A.dk);
      const int hi = spec_shell(kxsqr, A.f * (double)(H - 1), A.dk);
      if (b >= lo && b <= hi)
        for (int q = 0; q < 5; ++q) acc[q] += A.part[((size_t)q * A.nsh + b) * N + i];
    } else {
      if (b < H) acc[0] += A.party[(size_t)b * N + i];
      const int ikx = (int)((sqrt(kxsqr) / A.dk) + 0.5);
      if (ikx == b) acc[1] += A.partx[i];
    }
  }
  for (int q = 0; q < 5; ++q) red[q][tid] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < 5; ++q) red[q][tid] += red[q][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    if (shell) {
      for (int q = 0; q < 5; ++q) out[q * A.nsh + b] = red[q][0];
    } else if (b < A.nxy) {
      outy[b] = red[0][0];
      outx[b] = red[1][0];
    }
  }
}

}  // namespace chns

// Fused pass kernels of one ETD2RK stage (CHNS-2D, FP64, gfx950).
//
// Layouts in HBM (N = Nx = Ny, Ny2 = N/2+1):
//   spectral state / saved N_n : column-major  S[j*N + i]   (j = ky index, i = kx index)
//   row-pass inputs  (H*)      : 2-row tiled [N/2][ldJ][2], j < Jt (dealiased columns)
//   row-pass outputs (Mh*, G*) : 4-row tiled [N/4][ldF][4], j < Ny2
// Column kernels (x direction) read the state contiguously and touch the
// row-major arrays in chunks of C consecutive columns; row kernels (y
// direction) are contiguous everywhere.
//
// Passes of one stage (reference functions they replace):
//   A  colA : T(c_i) -> column iFFT -> Hc_i               (chem_potential.c:27-32, nonlin.c:282-284)
//   B  rowB : row c2r -> c -> muF (ternary) / P(c) (binary) -> row r2c -> Mh
//             (chem_potential.c:31-91, surface_tension.c:20-35, nonlin.c:217-230)
//   C  colC : column FFT of Mh -> mu = T(M + a k^2 c) -> column iFFT; column
//             iFFT of T(i ky psi), T(-i kx psi), T(i kx w), T(i ky w), T(i kx c_i)
//             (velocity.c, nonlin.c:30-48, surface_tension.c:116-136)
//   D  rowD : row c2r of all dealiased fields, products u.grad w, A_i grad c_i,
//             u.grad c_i, row r2c; W1 = -F(u.grad w) - i ky F(Fx)
//             (nonlin.c:56-92,266-348, surface_tension.c:138-220)
//   E  colE : column FFT, N_w = W1 + i kx Fy, N_c = -adv - (lam/g) k^2 M,
//             ETD2RK update with host-built coefficients (etd2rk.c:81-212),
//             then the next stage's pass A for the dealiased columns.
#pragma once
#include <type_traits>

#include "chns_fft.hpp"

namespace chns {

enum { MODEL_NS = 0, MODEL_BINARY = 1, MODEL_TERNARY = 2 };

// exact (uncontracted) ksqr = kx*kx + ky*ky, as the host computes it
__device__ __forceinline__ double ksq(double kx, double ky) {
  return __dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky));
}

template <int LOG2N> __device__ __forceinline__ int absidx(int i) {
  constexpr int N = 1 << LOG2N;
  return i <= N / 2 ? i : N - i;
}
template <int LOG2N> __device__ __forceinline__ double kxof(int i, double f) {
  constexpr int N = 1 << LOG2N;
  return f * (double)(i <= N / 2 ? i : i - N);   // functions.c:19-25
}

// columns per column-kernel workgroup
// PP = points per thread: 16 (radix-16 stages, two LDS exchanges at N = 2048)
// suits kernels that hold one column array across a transform; 8 keeps the
// registers of kernels holding two.  The workgroup holds C columns whose LDS
// slots stay within 150 KB (one workgroup per CU at N = 2048) or 80 KB.
template <int LOG2N, int PP = 8> struct ColCfg {
  static constexpr int N = 1 << LOG2N;
  static constexpr int P = (PP == 16 && LOG2N >= 8) ? 16 : 8;
  static constexpr int T = N / P;
  static constexpr int C0 = (256 / T) > 4 ? (256 / T) : 4;
  // N = 2048 with 8 points: 2 columns x 256 threads keeps a 256-VGPR budget
  static constexpr int C = (LOG2N >= 11 && P == 8) ? 2 : C0;
  static constexpr int NT = C * T;
};
// rows per row-kernel workgroup; RV = 1 (a schedule of the ensemble
// autotuner) halves them (at N = 2048 a row workgroup holds one row either way)
template <int LOG2N, int RV = 0> struct RowCfg {
  static constexpr int N = 1 << LOG2N;
  static constexpr int P = 8;
  static constexpr int T = N / P;
  static constexpr int R0 = (256 / T) > 1 ? (256 / T) : 1;
  static constexpr int R1 = (R0 >> RV) > 1 ? (R0 >> RV) : 1;
  static constexpr int R = R1 < N ? R1 : N;
  static constexpr int NT = R * T;
};

// Tiled layouts of the row-pass arrays: element (x, j) of a logical [N][ld]
// array lives at ((x/TR)*ld + j)*TR + x%TR.  A column workgroup (2 columns at
// N = 2048) then reads 32*TR contiguous bytes per row group: TF = 4 rows make
// that a whole 128-byte line for the row-pass outputs (Mh*, G*), which the
// column kernels read; TJ = 2 rows for the row-pass inputs (H*), which the
// column kernels write and the row kernels read.
constexpr int TJ = 2;
constexpr int TF = 4;
template <int TR>
__device__ __forceinline__ size_t tix(int x, int j, int ld) {
  return ((size_t)(x / TR) * ld + j) * TR + (x % TR);
}

// Row handled by lane-row r of this workgroup.  Workgroups holding fewer than
// TF rows are grouped so that the TF rows of one output tile go to blocks
// b, b+8, b+16, ... which sit on the same XCD (blocks are dealt to the 8 XCDs
// round-robin) and start together: their partial-line writes of the outputs
// merge in that L2, and the two rows of an input pair read one line.
template <int LOG2N, int RV = 0> __device__ __forceinline__ int row_of_block(int b, int r) {
  constexpr int R = RowCfg<LOG2N, RV>::R;
  if constexpr (R >= TF) {
    return b * R + r;
  } else {
    constexpr int G = TF / R;   // workgroups per output tile
    const int xcd = b & 7, idx = b >> 3;
    return TF * (((idx / G) << 3) + xcd) + (idx % G) * R + r;
  }
}
template <int LOG2N, int RV = 0> __device__ __forceinline__ int row_of(int r) {
  return row_of_block<LOG2N, RV>(blockIdx.x, r);
}

// Kernel arguments of the pass kernels: one instance passed by value, or a
// batch of same-size instances whose buffers and parameters sit in device
// arrays indexed by blockIdx.z (ensembles).  The pass bodies only see B, P.
struct Inst {
  Bufs B;
  KParams P;
};
struct InstBatch {
  const Bufs* __restrict__ B;
  const KParams* __restrict__ P;
};
__device__ __forceinline__ const Bufs& inst_bufs(const Inst& a) { return a.B; }
__device__ __forceinline__ const KParams& inst_kp(const Inst& a) { return a.P; }
__device__ __forceinline__ const Bufs& inst_bufs(const InstBatch& a) { return a.B[blockIdx.z]; }
__device__ __forceinline__ const KParams& inst_kp(const InstBatch& a) { return a.P[blockIdx.z]; }

// Work item (tile, group) of a column kernel launched on a 1-D grid of
// ceil(ntiles*ngroups/16)*16 blocks.  Items 2q and 2q+1 (neighbouring tiles)
// go to blocks b and b+8, which sit on the same XCD and start together, so
// the two halves of each 128-byte line of the 2-row tiled arrays they write
// meet in that L2.
// Returns false for padding blocks (the whole workgroup exits at once).
__device__ __forceinline__ bool col_item(int ntiles, int ngroups, int* tile, int* group) {
  const int b = blockIdx.x, xcd = b & 7, idx = b >> 3;
  const int item = 2 * (((idx >> 1) << 3) + xcd) + (idx & 1);
  if (item >= ntiles * ngroups) return false;
  *tile = item % ntiles;
  *group = item / ntiles;
  return true;
}

// Stage twiddles of the column kernels' transforms come from a per-workgroup
// LDS table (TwLds: 5.1 KiB at N = 2048 with 8 points per thread, beside the
// 72 KiB exchange array, still two workgroups per CU).  From the global table
// each transform would load ten 16-byte twiddles per thread, more vector-L1
// traffic than its column.  Declares LT (true) and twp, the table the
// transforms take, filled from the global table GTW; the fill is ordered
// before the first transform's first twiddle read by that transform's opening
// barrier.
#define COL_TWIDDLES(GTW)                                                         \
  using TL = TwLds<LOG2N, CF::P>;                                                 \
  constexpr bool LT = true;                                                       \
  static_assert(TL::size <= CF::NT, "one table entry per thread");                \
  __shared__ cd twl[TL::size];                                                    \
  const cd* twp = twl;                                                            \
  if ((int)threadIdx.x < TL::size) twl[threadIdx.x] = TL::entry(GTW, threadIdx.x)

// ---------------------------------------------------------------- column I/O
template <int LOG2N, int P, int TR>
__device__ __forceinline__ void col_load_rm(const cd* __restrict__ A, int ld, int j, bool ok, int t, cd* v) {
  constexpr int T = (1 << LOG2N) / P;
#pragma unroll
  for (int m = 0; m < P; ++m) v[m] = ok ? A[tix<TR>(t + m * T, j, ld)] : mk(0.0, 0.0);
}
template <int LOG2N, int P, int TR>
__device__ __forceinline__ void col_store_rm(cd* __restrict__ A, int ld, int j, bool ok, int t, const cd* v) {
  constexpr int T = (1 << LOG2N) / P;
  if (ok) {
#pragma unroll
    for (int m = 0; m < P; ++m) A[tix<TR>(t + m * T, j, ld)] = v[m];
  }
}
// masked load of a state column: modes with |i| > a are truncated (functions.c:75-78)
template <int LOG2N, int P>
__device__ __forceinline__ void col_load_masked(const cd* __restrict__ S, int j, int a, int t, cd* v) {
  constexpr int N = 1 << LOG2N;
  constexpr int T = N / P;
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int i = t + m * T;
    v[m] = (absidx<LOG2N>(i) <= a) ? S[(size_t)j * N + i] : mk(0.0, 0.0);
  }
}

// SMOOTH_FILTER form of truncation(): every mode times e twice, e =
// exp(-36 pp^18) from the host table at (j, |i|) (functions.c:70-73)
__device__ __forceinline__ cd filt2(cd z, double e) { return cscale(cscale(z, e), e); }
template <int LOG2N>
__device__ __forceinline__ double filt_at(const double* __restrict__ ft, int j, int i) {
  constexpr int N = 1 << LOG2N;
  return ft[(size_t)j * (N / 2 + 1) + absidx<LOG2N>(i)];
}
// truncation() of a state column on load: the sharp cut (a = largest kept
// |i|), or with FILT the smooth filter (a < 0 still marks a column outside)
template <int LOG2N, int P, bool FILT>
__device__ __forceinline__ void col_load_trunc(const cd* __restrict__ S, const double* __restrict__ ft, int j, int a,
                                               int t, cd* v) {
  if constexpr (FILT) {
    constexpr int N = 1 << LOG2N;
    constexpr int T = N / P;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int i = t + m * T;
      v[m] = a >= 0 ? filt2(S[(size_t)j * N + i], filt_at<LOG2N>(ft, j, i)) : mk(0.0, 0.0);
    }
  } else {
    col_load_masked<LOG2N, P>(S, j, a, t, v);
  }
}

// ---------------------------------------------------------------- row I/O
// Assemble Z = X + iY (length N, Hermitian halves) from two dealiased
// half-spectra rows so one complex c2r yields two real rows.  Only real parts
// are used at ky = 0 and ky = N/2, as a c2r does.  YIK: Y is i*ky*Yrow.
// Row entries X[st*k], Y[st*k] for k < Jt (Y may be null: zero).
template <int LOG2N, bool YIK>
__device__ __forceinline__ void assemble_pair(const cd* X, const cd* Y, int st, int Jt, double f, int t, cd* v) {
  constexpr int N = 1 << LOG2N;
  constexpr int T = N / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int n = t + m * T;
    const int k = n <= N / 2 ? n : N - n;
    cd a = mk(0.0, 0.0), b = mk(0.0, 0.0);
    if (k < Jt) {
      a = X[st * k];
      if (Y) b = Y[st * k];
      if (YIK) b = ctimes_ik(b, f * (double)k);
    }
    if (n == 0 || n == N / 2) v[m] = mk(a.x, b.x);
    else if (n < N / 2) v[m] = mk(a.x - b.y, a.y + b.x);
    else v[m] = mk(a.x + b.y, b.x - a.y);
  }
}

// Pair from the TJ-row tiled arrays in global memory (row element k at +TJ*k).
template <int LOG2N, bool YIK>
__device__ __forceinline__ void row_load_pair(const cd* __restrict__ X, const cd* __restrict__ Y, int Jt,
                                              double f, int t, cd* v) {
  assemble_pair<LOG2N, YIK>(X, Y, TJ, Jt, f, t, v);
}

// One DMA copy: 16 bytes from src of each active lane l straight into LDS at
// byte address lds + 16 l (gfx950 global_load_lds_dwordx4; lds, from M0, is
// wave-uniform).  Written as asm so that the compiler's wait insertion does
// not see the copy: it would make every LDS read wait for all copies in
// flight (vmcnt(0)), as it cannot tell the staging buffers apart.  Readers
// wait for the copies themselves (RowStager); the compiler's own waits only
// count fewer operations than are outstanding, so they wait at least as long
// as they must.  M0 is restored.
__device__ __forceinline__ void dma_copy16(const cd* src, unsigned lds) {
  unsigned saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, off\n\t"
      "s_mov_b32 m0, %0"
      : "=&s"(saved)
      : "v"(src), "s"(lds)
      : "memory");
}

// Streams the dealiased rows of the next field pair into LDS while the
// current pair is transformed (one row per workgroup, T threads): lane l of a
// wave copies one 16-byte entry into slot base + l of the staging area.
// JM: the longest row, N/4 + 1 kept modes (N/2 + 1 with the smooth filter).
template <int LOG2N, int JM = (1 << LOG2N) / 4 + 1>
struct RowStager {
  static constexpr int N = 1 << LOG2N, T = N / 8, JMAX = JM;
  cd* stg;   // 2*JMAX slots: X in [0, Jt), Y in [Jt, 2 Jt)
  int Jt;
  template <bool HASY>
  __device__ __forceinline__ void issue(const cd* X, const cd* Y) const {
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int nslot = HASY ? 2 * Jt : Jt;
    for (int base = wave * 64; base < nslot; base += T) {
      const int s = base + lane;
      if (s < nslot) {
        const cd* src = X + TJ * s;
        if constexpr (HASY) {
          if (s >= Jt) src = Y + TJ * (s - Jt);
        }
        dma_copy16(src, (unsigned)(size_t)(__attribute__((address_space(3))) cd*)(stg + base));
      }
    }
  }
  // this thread's copies have landed (vmcnt(0)), then every wave's
  __device__ __forceinline__ void wait() const {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
  }
  // Loads one issue makes per wave at Jt = JMAX: ceil((nslot - 64 w) / T)
  // for wave w, at least min_loads(nslot) (the last wave).
  static constexpr int min_loads(int nslot) { return (nslot - T + 64 + T - 1) / T; }
  static constexpr int vmcnt_imm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x0F70; }
  // Copies of this stager have landed, while those of the one other stager
  // issued after it (HASY_NEWER) may still be in flight; then every wave's.
  // Vector memory operations complete in issue order, so waiting until at
  // most n are outstanding, n no more than the newer issue made, waits for
  // every one issued before that issue (later loads or stores only make it
  // wait longer).
  template <bool HASY_NEWER>
  __device__ __forceinline__ void wait_older() const {
    if (Jt == JMAX) {
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(min_loads(HASY_NEWER ? 2 * JMAX : JMAX)));
    } else {
      __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    lds_sync();
  }
  template <bool YIK, bool HASY>
  __device__ __forceinline__ void assemble(double f, int t, cd* v) const {
    assemble_pair<LOG2N, YIK>(stg, HASY ? stg + Jt : nullptr, 1, Jt, f, t, v);
  }
};

// Split the forward transform Z of a + ib into A = F(a), B = F(b) for
// n = t + m*T <= N/2 and hand them to op(n, A, B).  Uses lds (LdsSize entries).
// In two halves around the barrier: row_split_put writes Z to lds,
// row_split_take (after a workgroup barrier) reads the pairs.
template <int LOG2N, int S>
__device__ __forceinline__ void row_split_put(const cd* v, cd* lds, int t) {
  constexpr int T = (1 << LOG2N) / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) lds[lp(t + m * T) * S] = v[m];
}
template <int LOG2N, int S, class Op>
__device__ __forceinline__ void row_split_take(const cd* lds, int t, Op op) {
  constexpr int N = 1 << LOG2N;
  constexpr int T = N / 8;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const int n = t + m * T;
    if (n <= N / 2) {
      cd A, B;
      split_pair(lds[lp(n) * S], lds[lp((N - n) & (N - 1)) * S], &A, &B);
      op(n, A, B);
    }
  }
}
template <int LOG2N, int S, class Op>
__device__ __forceinline__ void row_split(const cd* v, cd* lds, int t, Op op) {
  row_split_put<LOG2N, S>(v, lds, t);
  lds_sync();
  row_split_take<LOG2N, S>(lds, t, op);
}

// Forward row transform of the pair a + ib in v (real-space layout of rfft)
// and its split, op(n, A, B) for n <= N/2: fused into the last stage of the
// 4-wave transform, the transform then row_split otherwise.  Tw: RowTw or the
// twiddle table.
// SYNC = false skips the opening barrier of the fused 4-wave transform (see
// fft_w4_fwd_split); the other transforms always synchronise.
template <int LOG2N, int R, bool SYNC = true, class Tw, class Op>
__device__ __forceinline__ void rfft_split(cd* v, cd* lds, const Tw& tws, int t, Op op) {
  if constexpr (LOG2N == 11 && R == 1) {
    if constexpr (std::is_same<Tw, RowTw<LOG2N, R>>::value) {
      fft_w4_fwd_split<true, SYNC>(v, lds, tws.w, t, op);
    } else {
      W4Tw<false> W;
      W.init(tws, t);
      fft_w4_fwd_split(v, lds, W, t, op);
    }
  } else {
    rfft<LOG2N, false, R>(v, lds, tws, t);
    row_split<LOG2N, R>(v, lds, t, op);
  }
}

// ============================================================ pass A
// grid: (ceil(Jt/C), nphase)
template <int LOG2N, class Args, bool FILT = false>
__global__ __launch_bounds__((ColCfg<LOG2N, 16>::NT)) void k_colA(Args args) {
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  using CF = ColCfg<LOG2N, 16>;
  constexpr int N = CF::N;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  const int j = blockIdx.x * CF::C + c;
  const bool ok = j < P.Jt;
  cd* lds = lds_all + c;
  COL_TWIDDLES(B.tw);   // as pass E's fused pass A: the same values bit for bit below 256^2
  const cd* S = blockIdx.y == 0 ? B.C1 : B.C2;
  cd* H = blockIdx.y == 0 ? B.Hc1 : B.Hc2;
  cd v[CF::P];
  col_load_trunc<LOG2N, CF::P, FILT>(S, B.filt, j, ok ? B.amask[j] : -1, t, v);
  fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
  col_store_rm<LOG2N, CF::P, TJ>(H, P.ldJ, j, ok, t, v);
}

// ============================================================ pass B
// The pointwise part of pass B at one point: muF_1, muF_2 times fac
// (chem_potential.c:40-86) for the ternary model, P(c) = c(1-c)(1-2c)
// (surface_tension.c:126-127, nonlin.c:225) for the binary one.
template <int MODEL>
__device__ __forceinline__ cd muf_point(double c1, double c2, const KParams& P) {
  // f(c) = c (1 - c)(1 - 2c), evaluated as c + c^2 (2c - 3)
  auto f = [](double c) { return fma(c * c, fma(2.0, c, -3.0), c); };
  if constexpr (MODEL == MODEL_TERNARY) {
    const double c3 = (1.0 - c1 - c2);
    // X6 c1 c2 c3 times c_j c_k for G_i
    const double c23 = c2 * c3;
    const double q = P.X6 * (c1 * c23);
    const double G1 = fma(P.g1, f(c1), q * c23);
    const double G2 = fma(P.g2, f(c2), q * (c1 * c3));
    const double G3 = fma(P.g3, f(c3), q * (c1 * c2));
    // fac enters through the inverse coefficients (invariant across points)
    const double g1f = P.g1i * P.fac, g2f = P.g2i * P.fac, g3f = P.g3i * P.fac;
    const double d12 = G1 - G2;
    const double m1 = fma(d12, g2f, (G1 - G3) * g3f);
    const double m2 = fma(-d12, g1f, (G2 - G3) * g3f);
    return mk(m1, m2);
  } else {
    return mk(f(c1), 0.0);
  }
}

// grid: N / R
template <int LOG2N, int MODEL, class Args, bool FILT = false, int RV = 0>
__global__ __launch_bounds__((RowCfg<LOG2N, RV>::NT)) void k_rowB(Args args) {
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  using RC = RowCfg<LOG2N, RV>;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;   // rows interleaved across lanes
  const int x = row_of<LOG2N, RV>(r);
  cd* lds = lds_all + r;
  cd v[8];
  RowTw<LOG2N, RC::R> tws;
  tws.init(B.tw, t);
  const cd* h1 = B.Hc1 + tix<TJ>(x, 0, P.ldJ);
  const cd* h2 = (MODEL == MODEL_TERNARY) ? B.Hc2 + tix<TJ>(x, 0, P.ldJ) : nullptr;
  row_load_pair<LOG2N, false>(h1, h2, P.Jt, P.f, t, v);
  rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = muf_point<MODEL>(v[m].x * P.scale, v[m].y * P.scale, P);
  cd* o1 = B.Mh1 + tix<TF>(x, 0, P.ldF);
  cd* o2 = B.Mh2 + tix<TF>(x, 0, P.ldF);
  rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd a, cd b) {
    o1[TF * n] = a;
    if (MODEL == MODEL_TERNARY) o2[TF * n] = b;
  });
}

// Pass B at N = 2048 with the pruned N/4 cut, one row per workgroup at a
// time: workgroup b takes the rows of logical blocks b, b + G, b + 2G, ...
// (G = gridDim.x, a multiple of 32 so the rows of one output tile still run
// together on one XCD) and issues the next row's loads before transforming
// the current one, so they land while it computes instead of every row
// waiting for its own (a grid of one row per workgroup runs in rounds whose
// loads, transforms and stores do not overlap).  At 3 waves per SIMD (the
// next row's entries take 32 registers) G = 768 fills every CU once.
// With Jt <= N/4 + 1 a thread's nonzero row entries are those of its points
// m = 0, 1, 6, 7 (k = t, 256 + t, 512 - t, 256 - t), and thread 0's point
// m = 2 (n = 512), whose entry k = 512 is that of its point m = 6.
constexpr int kRowPipeM[4] = {0, 1, 6, 7};
struct RowRaw {
  cd a[4], b[4];
};
template <int LOG2N>
__device__ __forceinline__ void row_raw_load(const cd* __restrict__ X, const cd* __restrict__ Y, int Jt, int t,
                                             RowRaw& r) {
  constexpr int N = 1 << LOG2N, T = N / 8;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int n = t + kRowPipeM[q] * T;
    const int k = n <= N / 2 ? n : N - n;
    r.a[q] = k < Jt ? X[TJ * k] : mk(0.0, 0.0);
    r.b[q] = (Y && k < Jt) ? Y[TJ * k] : mk(0.0, 0.0);
  }
}
// Z = X + iY at the thread's points (assemble_pair on the raw entries)
template <int LOG2N>
__device__ __forceinline__ void row_raw_assemble(const RowRaw& r, int t, cd* v) {
  constexpr int N = 1 << LOG2N, T = N / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = mk(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = kRowPipeM[q];
    const int n = t + m * T;
    const cd a = r.a[q], b = r.b[q];
    if (n == 0) v[m] = mk(a.x, b.x);
    else if (n < N / 2) v[m] = mk(a.x - b.y, a.y + b.x);
    else v[m] = mk(a.x + b.y, b.x - a.y);
  }
  if (t == 0) v[2] = mk(r.a[2].x - r.b[2].y, r.a[2].y + r.b[2].x);   // n = 512 < N/2
}
template <int LOG2N, int MODEL, class Args>
__global__ __launch_bounds__(256, 3) void k_rowB_pipe(Args args) {
  static_assert(LOG2N == 11, "one row per 256-thread workgroup");
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  constexpr int N = 1 << LOG2N;
  __shared__ cd lds[RowLds<LOG2N, 1>::value];
  const int t = threadIdx.x;
  constexpr bool TER = MODEL == MODEL_TERNARY;
  RowTw<LOG2N, 1> tws;
  tws.init(B.tw, t);
  int L = blockIdx.x;
  const int G = gridDim.x;
  RowRaw raw;
  {
    const int x = row_of_block<LOG2N, 0>(L, 0);
    row_raw_load<LOG2N>(B.Hc1 + tix<TJ>(x, 0, P.ldJ), TER ? B.Hc2 + tix<TJ>(x, 0, P.ldJ) : nullptr, P.Jt, t, raw);
  }
  for (; L < N; L += G) {
    const int x = row_of_block<LOG2N, 0>(L, 0);
    cd v[8];
    row_raw_assemble<LOG2N>(raw, t, v);
    if (L + G < N) {
      const int xn = row_of_block<LOG2N, 0>(L + G, 0);
      row_raw_load<LOG2N>(B.Hc1 + tix<TJ>(xn, 0, P.ldJ), TER ? B.Hc2 + tix<TJ>(xn, 0, P.ldJ) : nullptr, P.Jt, t,
                          raw);
    }
    rfft<LOG2N, true, 1, true>(v, lds, tws, t);
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = muf_point<MODEL>(v[m].x * P.scale, v[m].y * P.scale, P);
    cd* o1 = B.Mh1 + tix<TF>(x, 0, P.ldF);
    cd* o2 = B.Mh2 + tix<TF>(x, 0, P.ldF);
    rfft_split<LOG2N, 1>(v, lds, tws, t, [&](int n, cd a, cd b) {
      o1[TF * n] = a;
      if (TER) o2[TF * n] = b;
    });
  }
}

// ============================================================ pass C
// grid: 1-D over (tile, group) items (col_item): ternary 6 groups, binary 4, NS 2.
// Groups are dispatched in order, the one-transform groups (d/dx c_i, and
// T(w) with CONS) last, so they fill the final round of workgroups.
// CONS: pass D takes the vorticity advection in conservative form, so the
// vorticity group writes T(w) (Hw) instead of T(i kx w), T(i ky w).
template <int LOG2N, int MODEL, class Args, bool FILT = false, int PP = 8, bool CONS = false>
__global__ __launch_bounds__((ColCfg<LOG2N, PP>::NT)) void k_colC(Args args) {
  static_assert(!CONS || (!FILT && MODEL != MODEL_NS), "conservative form: pruned binary/ternary passes only");
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  using CF = ColCfg<LOG2N, PP>;
  constexpr int N = CF::N, T = CF::T;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  constexpr int NG = MODEL == MODEL_TERNARY ? 6 : (MODEL == MODEL_BINARY ? 4 : 2);
  int tile, g;
  if (!col_item((P.Jt + CF::C - 1) / CF::C, NG, &tile, &g)) return;
  const int j = tile * CF::C + c;
  const bool ok = j < P.Jt;
  const int a = ok ? B.amask[j] : -1;
  cd* lds = lds_all + c;
  const double ky = P.f * (double)j;
  COL_TWIDDLES(B.tw);
  if (MODEL == MODEL_NS) g += 2;                 // NS: only the velocity groups
  if (MODEL == MODEL_BINARY && g >= 1) g += 1;   // binary: one mu group
  cd v[CF::P], s[CF::P];
  if (g <= 1) {
    // mu_i = T(M_i + 0.75 g_i eps k^2 c_i)  (surface_tension.c:116-128)
    // binary: mu = T(1.5 sigma eps k^2 c + 24 sigma/eps M)  (surface_tension.c:139-142,166)
    col_load_rm<LOG2N, CF::P, TF>(g == 0 ? B.Mh1 : B.Mh2, P.ldF, j, ok, t, v);
    fft<LOG2N, CF::P, false, CF::C, false, LT>(v, lds, twp, t);
    col_load_masked<LOG2N, CF::P>(g == 0 ? B.C1 : B.C2, j, a, t, s);
    const double mc = g == 0 ? P.mc1 : P.mc2;
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const int i = t + m * T;
      const double ks = ksq(kxof<LOG2N>(i, P.f), ky);
      if (absidx<LOG2N>(i) <= a) {
        if (MODEL == MODEL_TERNARY) v[m] = cadd(v[m], cscale(s[m], mc * ks));
        else v[m] = cadd(cscale(s[m], P.mcb * ks), cscale(v[m], P.s24));
        if constexpr (FILT) v[m] = filt2(v[m], filt_at<LOG2N>(B.filt, j, i));
      } else {
        v[m] = mk(0.0, 0.0);
      }
    }
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(g == 0 ? B.Hmu1 : B.Hmu2, P.ldJ, j, ok, t, v);
  } else if (g == 2) {
    // psi = w/k^2; u = i ky psi, v = -i kx psi (velocity.c:20-39)
    col_load_trunc<LOG2N, CF::P, FILT>(B.W, B.filt, j, a, t, s);
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const int i = t + m * T;
      const double ks = ksq(kxof<LOG2N>(i, P.f), ky);
      s[m] = ks > 1e-12 ? mk(s[m].x / ks, s[m].y / ks) : mk(0.0, 0.0);
      v[m] = ctimes_ik(s[m], ky);
    }
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(B.Hu, P.ldJ, j, ok, t, v);
#pragma unroll
    for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], -kxof<LOG2N>(t + m * T, P.f));
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(B.Hv, P.ldJ, j, ok, t, v);
  } else if (CONS && g == 3) {
    // T(w) (nonlin.c:36-37 without the derivatives, which pass D's
    // conservative form takes in the products' transforms)
    col_load_trunc<LOG2N, CF::P, FILT>(B.W, B.filt, j, a, t, v);
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(B.Hw, P.ldJ, j, ok, t, v);
  } else if (g == 3) {
    // grad w: i kx w, i ky w (nonlin.c:36-37)
    col_load_trunc<LOG2N, CF::P, FILT>(B.W, B.filt, j, a, t, s);
    if (P.sponge) {
      // T(omega) itself for the sponge term (nonlin.c:50-54)
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = s[m];
      fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
      col_store_rm<LOG2N, CF::P, TJ>(B.Hw, P.ldJ, j, ok, t, v);
    }
#pragma unroll
    for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], kxof<LOG2N>(t + m * T, P.f));
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(B.Hwx, P.ldJ, j, ok, t, v);
#pragma unroll
    for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], ky);
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(B.Hwy, P.ldJ, j, ok, t, v);
  } else {
    // d/dx c_i (nonlin.c:272,314; surface_tension.c:126,178): g = 4 c1, 5 c2
    col_load_trunc<LOG2N, CF::P, FILT>(g == 4 ? B.C1 : B.C2, B.filt, j, a, t, s);
#pragma unroll
    for (int m = 0; m < CF::P; ++m) v[m] = ctimes_ik(s[m], kxof<LOG2N>(t + m * T, P.f));
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(g == 4 ? B.Hdx1 : B.Hdx2, P.ldJ, j, ok, t, v);
  }
}

// ============================================================ pass D
// grid: N / R
// CONS (binary and ternary with the pruned N/4 cut, no sponge): the
// vorticity advection in conservative form (see the CONS branch); pass C
// then writes Hw instead of Hwx, Hwy.
template <int LOG2N, int MODEL, class Args, bool FILT = false, int RV = 0, bool CONS = false>
__global__ __launch_bounds__((RowCfg<LOG2N, RV>::NT), 2) void k_rowD(Args args) {
  static_assert(!CONS || (!FILT && MODEL != MODEL_NS), "conservative form: pruned binary/ternary passes only");
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  using RC = RowCfg<LOG2N, RV>;
  constexpr int N = RC::N, T = RC::T;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;   // rows interleaved across lanes
  const int x = row_of<LOG2N, RV>(r);
  cd* lds = lds_all + r;
  const size_t ro = tix<TJ>(x, 0, P.ldJ);   // row bases in the tiled layouts: element k at +TJ*k
  const size_t fo = tix<TF>(x, 0, P.ldF);   // ... and at +TF*k
  const double sc = P.scale;
  cd v[8];
  RowTw<LOG2N, RC::R> tws;
  tws.template init<RC::R == 1>(B.tw, t);   // the stager follows at one row per workgroup
  if constexpr (MODEL == MODEL_NS) {
    double u[8], w[8];
    // u, v then u.grad w (nonlin.c:47-48,62), cnlz = -F(nlz) (nonlin.c:74,92)
    row_load_pair<LOG2N, false>(B.Hu + ro, B.Hv + ro, P.Jt, P.f, t, v);
    rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
#pragma unroll
    for (int m = 0; m < 8; ++m) { u[m] = v[m].x * sc; w[m] = v[m].y * sc; }
    row_load_pair<LOG2N, false>(B.Hwx + ro, B.Hwy + ro, P.Jt, P.f, t, v);
    rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
    if (P.vpm) {
      // nlz, then -u chi/eta, -v chi/eta whose curl is subtracted (nonlin.c:62-95);
      // + omega sponge(y)/eta with the sponge (nonlin.c:50-54,68-70,97-99)
      double nl[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) nl[m] = u[m] * (v[m].x * sc) + w[m] * (v[m].y * sc);
      if (P.sponge) {
        row_load_pair<LOG2N, false>(B.Hw + ro, nullptr, P.Jt, P.f, t, v);
        rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
#pragma unroll
        for (int m = 0; m < 8; ++m) nl[m] += (v[m].x * sc) * B.spg[rpos<LOG2N, RC::R>(t, m)];
      }
      const double ce = B.chi[x] * P.etainv;
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(-nl[m], -u[m] * ce);
      cd* o = B.G1 + fo;
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd A, cd Bv) {
        const cd iky = ctimes_ik(Bv, P.f * (double)n);
        o[TF * n] = mk(A.x - iky.x, A.y - iky.y);
      });
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(-w[m] * ce, 0.0);
      cd* o2 = B.G2 + fo;
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd A, cd) { o2[TF * n] = A; });
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(u[m] * (v[m].x * sc) + w[m] * (v[m].y * sc), 0.0);
      cd* o = B.G1 + fo;
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd A, cd) { o[TF * n] = mk(-A.x, -A.y); });
    }
  } else if constexpr (CONS) {
    // For a divergence-free u = (u, v), u.grad w = d/dx(u w) + d/dy(v w), so
    // the vorticity right-hand side (nonlin.c:56-92 with the curl of the
    // surface-tension force, curl.c:16-31)
    //   -F(u.grad w) - i ky F(Fx) + i kx F(Fy) = i kx F(P) - i ky F(Q),
    //   P = Fy - u w,  Q = Fx + v w,
    // takes the row c2r of w itself instead of those of dw/dx and dw/dy, and
    // one paired forward transform (P, Q) instead of -u.grad w, Fx and Fy:
    // G1 = -i ky F(Q) (the ky factor of this row transform), G2 = F(P), and
    // pass E adds i kx F(G2) as before.  Both forms are the same function of
    // the dealiased state: every kept mode has |kx|, |ky| <= N/4, so the
    // grid products alias only onto kx = N/2 (from the modes (+-N/4, 0),
    // where u = 0 and d/dy = 0) and ky = N/2 (from (0, +-N/4), where v = 0
    // and d/dx = 0), and both forms vanish there; they differ by round-off.
    // Ternary (binary: mu and w share the first pair, c2 is absent):
    //   (mu1, mu2) -> A_i; (u, v) -> Q = -u chi/eta, P = -v chi/eta (walls);
    //   grad c1, grad c2 -> Q += A_i dc_i/dx, P += A_i dc_i/dy, adv_i;
    //   forward (adv1, adv2); w -> Q += v w, P -= u w; forward (P, Q).
    // Walls: A_i and adv_i carry the factor 1 - chi (surface_tension.c:
    // 83-85,151-154,202-205; nonlin.c:209-211,294-296,337-339).
    constexpr bool TER = MODEL == MODEL_TERNARY;
    constexpr bool STG = RC::R == 1;   // one row per workgroup: the field pairs are staged in LDS
    __shared__ cd stgA[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
    __shared__ cd stgB[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
    const RowStager<LOG2N> sa{stgA, P.Jt}, sb{stgB, P.Jt};
    const double om = P.vpm ? 1.0 - B.chi[x] : 1.0;
    const double ce = P.vpm ? B.chi[x] * P.etainv : 0.0;
    const cd* const hm2 = TER ? B.Hmu2 + ro : B.Hw + ro;
    double ux[8], uy[8], pp[8], qq[8], ad1[8], ad2[8];
    {
      double A1[8], A2[8];   // binary: A2 holds w
      if constexpr (STG) {
        sa.template issue<true>(B.Hmu1 + ro, hm2);
        sb.template issue<true>(B.Hu + ro, B.Hv + ro);
        sa.template wait_older<true>();
        sa.template assemble<false, true>(P.f, t, v);
        lds_sync();
        sa.template issue<true>(B.Hdx1 + ro, B.Hc1 + ro);
      } else {
        row_load_pair<LOG2N, false>(B.Hmu1 + ro, hm2, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, true, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double m1 = v[m].x * sc, m2 = v[m].y * sc;
        if (TER) {
          A1[m] = (P.A11 * m1 + P.A12 * m2) * om;   // surface_tension.c:144
          A2[m] = (P.A21 * m1 + P.A22 * m2) * om;   // surface_tension.c:195
        } else {
          A1[m] = m1 * om;                          // binary mu (surface_tension.c:183-184)
          A2[m] = m2;                               // w
        }
      }
      if constexpr (STG) {
        sb.template wait_older<true>();
        sb.template assemble<false, true>(P.f, t, v);
        if (TER) {
          lds_sync();
          sb.template issue<true>(B.Hdx2 + ro, B.Hc2 + ro);
        }
      } else {
        row_load_pair<LOG2N, false>(B.Hu + ro, B.Hv + ro, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, true, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        ux[m] = v[m].x * sc;
        uy[m] = v[m].y * sc;
        if (TER) {
          qq[m] = -(ux[m] * ce);
          pp[m] = -(uy[m] * ce);
        } else {
          qq[m] = uy[m] * A2[m] - ux[m] * ce;
          pp[m] = -(ux[m] * A2[m]) - uy[m] * ce;
        }
      }
      if constexpr (STG) {
        if (TER) {
          sa.template wait_older<true>();
        } else {
          sa.wait();
        }
        sa.template assemble<true, true>(P.f, t, v);
        if (TER) {
          lds_sync();
          sa.template issue<false>(B.Hw + ro, nullptr);
        }
      } else {
        row_load_pair<LOG2N, true>(B.Hdx1 + ro, B.Hc1 + ro, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, true, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double dx = v[m].x * sc, dy = v[m].y * sc;
        qq[m] += dx * A1[m];
        pp[m] += dy * A1[m];
        ad1[m] = dx * ux[m] + dy * uy[m];
      }
      if constexpr (TER) {
        if constexpr (STG) {
          sb.template wait_older<false>();
          sb.template assemble<true, true>(P.f, t, v);
        } else {
          row_load_pair<LOG2N, true>(B.Hdx2 + ro, B.Hc2 + ro, P.Jt, P.f, t, v);
        }
        rfft<LOG2N, true, RC::R, true, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const double dx = v[m].x * sc, dy = v[m].y * sc;
          qq[m] += dx * A2[m];
          pp[m] += dy * A2[m];
          ad2[m] = dx * ux[m] + dy * uy[m];
        }
      }
    }
    if (P.vpm) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        ad1[m] *= om;
        if (TER) ad2[m] *= om;
      }
    }
    cd* o1 = B.G1 + fo;
    cd* o2 = B.G2 + fo;
    cd* o3 = B.G3 + fo;
    const auto put12 = [&](int n, cd A, cd Bq) {
      const cd iky = ctimes_ik(Bq, P.f * (double)n);
      o1[TF * n] = mk(-iky.x, -iky.y);
      o2[TF * n] = A;
    };
    if constexpr (TER) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(ad1[m], ad2[m]);
      cd* o4 = B.G4 + fo;
      const auto put34 = [&](int n, cd A, cd Bv) {
        o3[TF * n] = A;
        o4[TF * n] = Bv;
      };
      if constexpr (STG) {
        // wait for the w copy before the stores of this transform are
        // issued, so that the wait (vmcnt(0)) is not also for them
        sa.wait();
        rfft_split<LOG2N, RC::R, false>(v, lds, tws, t, put34);   // after the wait's barrier
        sa.template assemble<false, false>(P.f, t, v);
      } else {
        rfft_split<LOG2N, RC::R>(v, lds, tws, t, put34);
        row_load_pair<LOG2N, false>(B.Hw + ro, nullptr, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, true>(v, lds, tws, t);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double w = v[m].x * sc;
        v[m] = mk(pp[m] - ux[m] * w, qq[m] + uy[m] * w);
      }
      // no other wave touches this wave's lds quarter after the w inverse
      // transform's exchange
      rfft_split<LOG2N, RC::R, !STG>(v, lds, tws, t, put12);
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(pp[m], qq[m]);
      rfft_split<LOG2N, RC::R, !STG>(v, lds, tws, t, put12);   // after the grad c inverse transform
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(ad1[m], 0.0);
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd A, cd) { o3[TF * n] = A; });
    }
  } else {
    // Order: mu -> A_i, then velocity, then grad c_i (surface-tension force
    // F = sum_i A_i grad c_i, surface_tension.c:138-220, and advection u.grad c_i,
    // nonlin.c:266-348, accumulate as each pair lands), then grad w for u.grad w
    // (nonlin.c:56-62).  (Fy, u.grad c1) are transformed as soon as they are
    // complete; -u.grad w and Fx share the next forward transform, so
    // W1 = -F(u.grad w) - i ky F(Fx) (curl.c:26, ky part) comes out of its split.
    constexpr bool TER = MODEL == MODEL_TERNARY;
    // One row per workgroup (N = 2048): the field pairs are staged in LDS,
    // alternately in two buffers, so pair k+2 streams in while pairs k and
    // k+1 are transformed.
    constexpr bool STG = RC::R == 1 && !FILT;   // the stager holds dealiased rows
    __shared__ cd stgA[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
    __shared__ cd stgB[STG ? 2 * RowStager<LOG2N>::JMAX : 1];
    const RowStager<LOG2N> sa{stgA, P.Jt}, sb{stgB, P.Jt};
    // Smooth filter: every mode is kept, so the rows are twice as long and a
    // single buffer streams pair k+1 in while pair k is transformed (two
    // buffers would leave one workgroup per CU).
    constexpr bool STG1 = RC::R == 1 && FILT;
    constexpr int JF = RC::N / 2 + 1;
    __shared__ cd stgF[STG1 ? 2 * JF : 1];
    const RowStager<LOG2N, JF> s1{stgF, P.Jt};
    double u[8], w[8], fx[8], fy[8], ad1[8], ad2[8];
    {
      double A1[8], A2[8];
      if constexpr (STG) {
        sa.template issue<TER>(B.Hmu1 + ro, B.Hmu2 + ro);
        sb.template issue<true>(B.Hu + ro, B.Hv + ro);
        sa.template wait_older<true>();
        sa.template assemble<false, TER>(P.f, t, v);
        lds_sync();
        sa.template issue<true>(B.Hdx1 + ro, B.Hc1 + ro);
      } else if constexpr (STG1) {
        s1.template issue<TER>(B.Hmu1 + ro, B.Hmu2 + ro);
        s1.wait();
        s1.template assemble<false, TER>(P.f, t, v);
        lds_sync();
        s1.template issue<true>(B.Hu + ro, B.Hv + ro);
      } else {
        row_load_pair<LOG2N, false>(B.Hmu1 + ro, TER ? B.Hmu2 + ro : nullptr, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, !FILT, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double m1 = v[m].x * sc, m2 = v[m].y * sc;
        if (TER) {
          A1[m] = (P.A11 * m1 + P.A12 * m2);   // surface_tension.c:144
          A2[m] = (P.A21 * m1 + P.A22 * m2);   // surface_tension.c:195
        } else {
          A1[m] = m1;                          // binary mu (surface_tension.c:183-184)
        }
      }
      if constexpr (STG) {
        sb.template wait_older<true>();
        sb.template assemble<false, true>(P.f, t, v);
        lds_sync();
        if (TER) sb.template issue<true>(B.Hdx2 + ro, B.Hc2 + ro);
        else sb.template issue<true>(B.Hwx + ro, B.Hwy + ro);
      } else if constexpr (STG1) {
        s1.wait();
        s1.template assemble<false, true>(P.f, t, v);
        lds_sync();
        s1.template issue<true>(B.Hdx1 + ro, B.Hc1 + ro);
      } else {
        row_load_pair<LOG2N, false>(B.Hu + ro, B.Hv + ro, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, !FILT, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        u[m] = v[m].x * sc;
        w[m] = v[m].y * sc;
      }
      if constexpr (STG) {
        sa.template wait_older<true>();
        sa.template assemble<true, true>(P.f, t, v);
        if (TER) {
          lds_sync();
          sa.template issue<true>(B.Hwx + ro, B.Hwy + ro);
        }
      } else if constexpr (STG1) {
        s1.wait();
        s1.template assemble<true, true>(P.f, t, v);
        lds_sync();
        if (TER) s1.template issue<true>(B.Hdx2 + ro, B.Hc2 + ro);
        else s1.template issue<true>(B.Hwx + ro, B.Hwy + ro);
      } else {
        row_load_pair<LOG2N, true>(B.Hdx1 + ro, B.Hc1 + ro, P.Jt, P.f, t, v);
      }
      rfft<LOG2N, true, RC::R, !FILT, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double dx = v[m].x * sc, dy = v[m].y * sc;
        fx[m] = dx * A1[m];
        fy[m] = dy * A1[m];
        ad1[m] = dx * u[m] + dy * w[m];
      }
      if constexpr (TER) {
        if constexpr (STG) {
          sb.template wait_older<true>();
          sb.template assemble<true, true>(P.f, t, v);
        } else if constexpr (STG1) {
          s1.wait();
          s1.template assemble<true, true>(P.f, t, v);
          lds_sync();
          s1.template issue<true>(B.Hwx + ro, B.Hwy + ro);
        } else {
          row_load_pair<LOG2N, true>(B.Hdx2 + ro, B.Hc2 + ro, P.Jt, P.f, t, v);
        }
        rfft<LOG2N, true, RC::R, !FILT, !STG>(v, lds, tws, t);   // after a stager barrier
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const double dx = v[m].x * sc, dy = v[m].y * sc;
          fx[m] += dx * A2[m];
          fy[m] += dy * A2[m];
          ad2[m] = dx * u[m] + dy * w[m];
        }
      }
    }
    // walls: F and u.grad c_i times (1 - chi), F minus u chi/eta (surface_tension.c:
    // 83-85,151-154,202-205; nonlin.c:63-66,209-211,294-296,337-339)
    const double om = P.vpm ? 1.0 - B.chi[x] : 1.0;
    const double ce = P.vpm ? B.chi[x] * P.etainv : 0.0;
    if (P.vpm) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        fx[m] = fx[m] * om - u[m] * ce;
        fy[m] = fy[m] * om - w[m] * ce;
        ad1[m] *= om;
        if (TER) ad2[m] *= om;
      }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = mk(fy[m], ad1[m]);
    cd* o2 = B.G2 + fo;
    cd* o3 = B.G3 + fo;
    const auto put23 = [&](int n, cd A, cd Bv) {
      o2[TF * n] = A;
      o3[TF * n] = Bv;
    };
    if constexpr (STG) {
      // wait for the grad w pair before the stores of this transform are
      // issued, so that the wait (vmcnt(0)) is not also for them
      const RowStager<LOG2N>& sw = TER ? sa : sb;
      sw.wait();
      rfft_split<LOG2N, RC::R, false>(v, lds, tws, t, put23);   // after the wait's barrier
      sw.template assemble<false, true>(P.f, t, v);
    } else if constexpr (STG1) {
      s1.wait();
      rfft_split<LOG2N, RC::R, false>(v, lds, tws, t, put23);   // after the wait's barrier
      s1.template assemble<false, true>(P.f, t, v);
    } else {
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, put23);
      row_load_pair<LOG2N, false>(B.Hwx + ro, B.Hwy + ro, P.Jt, P.f, t, v);
    }
    rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
    if (P.sponge) {
      // + omega sponge(y)/eta in nlz (nonlin.c:50-54,68-70,97-99)
      double nl[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) nl[m] = u[m] * (v[m].x * sc) + w[m] * (v[m].y * sc);
      row_load_pair<LOG2N, false>(B.Hw + ro, nullptr, P.Jt, P.f, t, v);
      rfft<LOG2N, true, RC::R, !FILT>(v, lds, tws, t);
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(-(nl[m] + (v[m].x * sc) * B.spg[rpos<LOG2N, RC::R>(t, m)]), fx[m]);
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(-(u[m] * (v[m].x * sc) + w[m] * (v[m].y * sc)), fx[m]);
    }
    cd* o1 = B.G1 + fo;
    // no other wave touches this wave's lds quarter after the grad w inverse
    // transform's exchange (the sponge's is alike)
    rfft_split<LOG2N, RC::R, !STG>(v, lds, tws, t, [&](int n, cd A, cd Bv) {
      const cd iky = ctimes_ik(Bv, P.f * (double)n);
      o1[TF * n] = mk(A.x - iky.x, A.y - iky.y);
    });
    if constexpr (TER) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = mk(ad2[m], 0.0);
      cd* o4 = B.G4 + fo;
      rfft_split<LOG2N, RC::R>(v, lds, tws, t, [&](int n, cd A, cd) { o4[TF * n] = A; });
    }
  }
}

// ============================================================ pass E
// grid: 1-D over (tile, field) items (col_item): omega, then phase field(s).
template <int LOG2N, int MODEL, int STAGE, class Args, bool FILT = false, int PP = 8>
__global__ __launch_bounds__((ColCfg<LOG2N, PP>::NT)) void k_colE(Args args) {
  const Bufs& B = inst_bufs(args);
  const KParams& P = inst_kp(args);
  using CF = ColCfg<LOG2N, PP>;
  constexpr int N = CF::N, T = CF::T, H = N / 2 + 1;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  constexpr int NG = MODEL == MODEL_TERNARY ? 3 : (MODEL == MODEL_BINARY ? 2 : 1);
  int tile, g;
  // omega items first, then the phase-field items in pairs (c1 tile, c2 tile)
  // on one XCD at one time, so the second reads the coefficient tables from L2
  {
    const int ntl = (H + CF::C - 1) / CF::C, nw = (ntl + 1) & ~1;
    int item, dummy;
    if (!col_item(nw + (NG - 1) * ntl, 1, &item, &dummy)) return;
    if (item < nw) {
      if (item >= ntl) return;
      tile = item;
      g = 0;
    } else if constexpr (NG > 1) {
      const int r = item - nw;
      tile = r / (NG - 1);
      g = 1 + r % (NG - 1);
    } else {
      return;
    }
  }

  const int j = tile * CF::C + c;
  const bool ok = j < H;
  cd* lds = lds_all + c;
  const double ky = P.f * (double)j;
  COL_TWIDDLES(B.tw);
  cd v[CF::P], s[CF::P];
  const cd* Gm = g == 0 ? B.G1 : (g == 1 ? B.G3 : B.G4);
  col_load_rm<LOG2N, CF::P, TF>(Gm, P.ldF, j, ok, t, v);
  fft<LOG2N, CF::P, false, CF::C, false, LT>(v, lds, twp, t);
  if (g == 0) {
    if (MODEL != MODEL_NS || P.vpm) {
      // + i kx F(Fy): the curl of the surface-tension force (curl.c:26) and of
      // the wall penalisation (nonlin.c:76-80)
      col_load_rm<LOG2N, CF::P, TF>(B.G2, P.ldF, j, ok, t, s);
      fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = cadd(v[m], ctimes_ik(s[m], kxof<LOG2N>(t + m * T, P.f)));
    }
    if (P.grav && ok) {
      // buoyancy torque (nonlin.c:107-171), from k_gravity
#pragma unroll
      for (int m = 0; m < CF::P; ++m) v[m] = cadd(v[m], B.GR[(size_t)j * N + t + m * T]);
    }
    if (P.f_k > 0 && j == P.f_k && t == 0) v[0].y += P.force_im;   // Kolmogorov forcing at (0, f_k)
  } else {
    // N_c = -F(u.grad c) - (lam/g) k^2 M  (nonlin.c:369-370); binary: - lam*(24 s/e k^2 M) (:240,251)
    col_load_rm<LOG2N, CF::P, TF>(g == 1 ? B.Mh1 : B.Mh2, P.ldF, j, ok, t, s);
    fft<LOG2N, CF::P, false, CF::C, false, LT>(s, lds, twp, t);
    const double lg = g == 1 ? P.lg1 : P.lg2;
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const double ks = ksq(kxof<LOG2N>(t + m * T, P.f), ky);
      if (MODEL == MODEL_TERNARY) {
        v[m] = csub(mk(-v[m].x, -v[m].y), cscale(s[m], lg * ks));
      } else {
        v[m] = csub(mk(-v[m].x, -v[m].y), cscale(cscale(s[m], P.s24 * ks), P.lam));
      }
    }
  }
  // ETD2RK update (etd2rk.c:99,113,129-130 stage 1; :174,188,205-206 stage 2)
  const cd* S = g == 0 ? B.W : (g == 1 ? B.C1 : B.C2);
  cd* So = g == 0 ? B.Wo : (g == 1 ? B.C1o : B.C2o);
  cd* Nsv = g == 0 ? B.NW : (g == 1 ? B.NC1 : B.NC2);
  // Stage 1 keeps b = a - z2 N(u) instead of N(u), so stage 2 forms
  // u' = a + z2 (N(a) - N(u)) as b + z2 N(a) without reading a back.
  // fin: every updated mode of this thread is finite (re + im is finite
  // unless one part is inf or NaN, or both overflow together)
  bool fin = true;
  // Stage 1 in place: only the kept modes of a are ever read (pass C of stage
  // 2 reads T(a), pass A is fused below, stage 2 forms u' from b), so the
  // others are not written.  Written in full when the unpruned kernels read
  // every mode (FILT), for the buoyancy term (k_gravity reads the
  // untruncated state) and when writing out of place (the adjoint's replay
  // keeps a).
  const bool fullA = FILT || P.grav || So != S;
  const int keptA = fullA ? N : (j < P.Jt ? B.amask[j] : -1);
  if (ok) {
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const int i = t + m * T;
      const size_t idx = (size_t)j * N + i;
      const int ti = j * H + absidx<LOG2N>(i);
      const double z2 = (g == 0 ? B.tabw2 : B.tabc2)[ti];
      if constexpr (STAGE == 1) {
        const cd u0 = S[idx];
        const cd k1 = (g == 0 ? B.tabw1 : B.tabc1)[ti];
        const double e = k1.x, z = k1.y;
        const cd a = cadd(cscale(u0, e), cscale(v[m], z));
        if (absidx<LOG2N>(i) <= keptA) So[idx] = a;
        Nsv[idx] = csub(a, cscale(v[m], z2));
        s[m] = a;
        fin = fin && __builtin_isfinite(a.x + a.y);
      } else {
        const cd a = cadd(Nsv[idx], cscale(v[m], z2));
        So[idx] = a;
        s[m] = a;
        fin = fin && __builtin_isfinite(a.x + a.y);
      }
    }
  }
  if (!fin) *B.flag = 1;   // a plain store of the same value from any lane (read by chns_status)
  // next stage's pass A for the phase fields
  if (g >= 1 && tile * CF::C < P.Jt) {
    const bool okA = j < P.Jt;
    const int a = okA ? B.amask[j] : -1;
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const int i = t + m * T;
      if constexpr (FILT) v[m] = okA ? filt2(s[m], filt_at<LOG2N>(B.filt, j, i)) : mk(0.0, 0.0);
      else v[m] = (okA && absidx<LOG2N>(i) <= a) ? s[m] : mk(0.0, 0.0);
    }
    fft<LOG2N, CF::P, true, CF::C, !FILT, LT>(v, lds, twp, t);
    col_store_rm<LOG2N, CF::P, TJ>(g == 1 ? B.Hc1 : B.Hc2, P.ldJ, j, okA, t, v);
  }
}

// ============================================================ dfft + truncation of real fields
// Row r2c of packed real [N][N] fields a and b (b may be null), two rows per
// complex transform; outputs row-major [N][ldF] half-spectra.  grid: N / R
template <int LOG2N>
__global__ __launch_bounds__(RowCfg<LOG2N>::NT) void k_row_r2c(const double* __restrict__ a,
                                                               const double* __restrict__ b, cd* oa, cd* ob,
                                                               const cd* __restrict__ tw, int ldF) {
  using RC = RowCfg<LOG2N>;
  constexpr int N = RC::N, T = RC::T;
  __shared__ cd lds_all[RowLds<LOG2N, RC::R>::value];
  const int r = threadIdx.x % RC::R, t = threadIdx.x / RC::R;   // rows interleaved across lanes
  const int x = row_of<LOG2N>(r);
  cd* lds = lds_all + r;
  cd v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const size_t o = (size_t)x * N + rpos<LOG2N, RC::R>(t, m);
    v[m] = mk(a[o], b ? b[o] : 0.0);
  }
  cd* pa = oa + tix<TF>(x, 0, ldF);
  cd* pb = ob ? ob + tix<TF>(x, 0, ldF) : nullptr;
  rfft_split<LOG2N, RC::R>(v, lds, tw, t, [&](int n, cd A, cd Bv) {
    pa[TF * n] = A;
    if (pb) pb[TF * n] = Bv;
  });
}

// Column FFT of row-major half-transformed data into the column-major state,
// with truncation() applied when TRUNC (init.c:256-262,291-294).
// grid: (ceil(Ny2/C), 1)
template <int LOG2N, bool TRUNC>
__global__ __launch_bounds__((ColCfg<LOG2N, 16>::NT)) void k_col_fwd(const cd* __restrict__ in, cd* out,
                                                              const cd* __restrict__ tw,
                                                              const int* __restrict__ amask,
                                                              const double* __restrict__ ft, int ldF) {
  using CF = ColCfg<LOG2N, 16>;
  constexpr int N = CF::N, T = CF::T, H = N / 2 + 1;
  __shared__ cd lds_all[CF::C * LdsSize<LOG2N>::value];
  const int c = threadIdx.x % CF::C, t = threadIdx.x / CF::C;
  const int j = blockIdx.x * CF::C + c;
  const bool ok = j < H;
  cd* lds = lds_all + c;
  cd v[CF::P];
  col_load_rm<LOG2N, CF::P, TF>(in, ldF, j, ok, t, v);
  fft<LOG2N, CF::P, false, CF::C>(v, lds, tw, t);
  if (ok) {
    const int a = TRUNC ? amask[j] : N;
#pragma unroll
    for (int m = 0; m < CF::P; ++m) {
      const int i = t + m * T;
      const cd z = TRUNC && ft ? filt2(v[m], filt_at<LOG2N>(ft, j, i)) : v[m];   // SMOOTH_FILTER
      out[(size_t)j * N + i] = (absidx<LOG2N>(i) <= a) ? z : mk(0.0, 0.0);
    }
  }
}

// Buoyancy torque i kx F(-g (rho(c) - 1)) of nonlin.c:107-171.  rho is linear
// in the untruncated c, so the transform pair of the reference reduces to
// GR = i kx (gr1 c1 + gr2 c2) on the spectral state (the constant rho3 - 1
// only reaches kx = 0).  Runs before pass E, which updates the state in place.
// grid: ceil(Ny2*N / 256)
static __global__ void k_gravity(const cd* __restrict__ C1, const cd* __restrict__ C2, cd* GR, double gr1,
                                 double gr2, double f, int N) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * (N / 2 + 1)) return;
  const int i = (int)(e % N);
  const double kx = f * (double)(i <= N / 2 ? i : i - N);
  cd r = cscale(C1[e], gr1);
  if (C2) r = cadd(r, cscale(C2[e], gr2));
  GR[e] = ctimes_ik(r, kx);
}

// In-place truncation of a column-major spectral field (spectrum.c:24-26):
// the sharp cut, or with ft the smooth filter (functions.c:70-73).
// grid: ceil(Ny2*N / 256)
static __global__ void k_truncate(cd* S, const int* __restrict__ amask, const double* __restrict__ ft, int N, int H) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)N * H) return;
  const int j = (int)(e / N), i = (int)(e % N);
  const int ai = i <= N / 2 ? i : N - i;
  if (ft) S[e] = filt2(S[e], ft[(size_t)j * H + ai]);
  else if (ai > amask[j]) S[e] = mk(0.0, 0.0);
}

// out[c][r] = in[r][c] for a row-major [rows][cols] complex array: the state
// between the reference layout IDXC [Nx][Ny/2+1] and the column-major
// [Ny/2+1][Nx] the passes use.  32 x 32 tiles through LDS (one padding
// column, so the transposed reads hit distinct banks); both the loads and the
// stores of a 32-lane row are 512 contiguous bytes.
// grid: (ceil(cols/32), ceil(rows/32)), 256 threads
static __global__ __launch_bounds__(256) void k_transpose(const cd* __restrict__ in, cd* __restrict__ out, int rows,
                                                          int cols) {
  __shared__ cd tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int r = r0 + ty + k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + k][tx] = in[(size_t)r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int c = c0 + ty + k, r = r0 + tx;
    if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][ty + k];
  }
}

}  // namespace chns

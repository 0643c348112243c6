// FP64 Stockham FFT engine for gfx950, staged through LDS.
//
// One transform of length N = 2^LOG2N is carried by T = N/P threads (P = 8 or
// 16 points per thread); thread t holds the points t + m*T (m < P) in
// registers ("cyclic" distribution) both before and after the transform, so a
// c2r can feed a pointwise stage and an r2c without extra data movement.
//
// Stages are radix P (registers) with one smaller final radix when LOG2N is
// not a multiple of log2(P).  Between stages the values go through LDS: every
// thread writes its butterfly outputs to their Stockham positions, the
// workgroup synchronises, and every thread reads its cyclic set back.  The
// synchronisations are workgroup-wide, so all threads of a workgroup must call
// the same transforms in the same order (inputs of idle lanes are zero).
//
// Twiddles come from a table tw[n] = exp(-2*pi*i*n/N), n < N, built on the
// host in extended precision, or from a workgroup's LDS table of each
// stage's base twiddles filled from it (TwLds, the column kernels); the
// inverse uses the conjugates.  No normalisation is applied (the 1/(Nx*Ny)
// of the c2r is applied by callers).
#pragma once
#include "chns_types.hpp"

namespace chns {

__device__ __forceinline__ cd mk(double a, double b) { cd r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmulc(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cscale(cd a, double s) { return mk(a.x * s, a.y * s); }
// (i*k)*z with C99 semantics: (-k*Im z, k*Re z)
__device__ __forceinline__ cd ctimes_ik(cd z, double k) { return mk(-k * z.y, k * z.x); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ cd rot90(cd a) { return INV ? mk(-a.y, a.x) : mk(a.y, -a.x); }
// multiply by exp(-+ 2 pi i e/16) given (cos, sin) of 2 pi e/16
template <bool INV> __device__ __forceinline__ cd wmul(cd a, double c, double s) {
  return INV ? mk(a.x * c - a.y * s, a.x * s + a.y * c) : mk(a.x * c + a.y * s, a.y * c - a.x * s);
}

constexpr double kR = 0.70710678118654752440084436210485;   // cos(pi/4)
constexpr double kC8 = 0.92387953251128675612818318939679;  // cos(pi/8)
constexpr double kS8 = 0.38268343236508977172845998403040;  // sin(pi/8)

template <bool INV> __device__ __forceinline__ void bfly2(cd& a, cd& b) {
  cd t = a; a = cadd(t, b); b = csub(t, b);
}

template <bool INV> __device__ __forceinline__ void dft4(cd& a0, cd& a1, cd& a2, cd& a3) {
  cd t0 = cadd(a0, a2), t1 = csub(a0, a2);
  cd t2 = cadd(a1, a3), t3 = rot90<INV>(csub(a1, a3));
  a0 = cadd(t0, t2); a2 = csub(t0, t2);
  a1 = cadd(t1, t3); a3 = csub(t1, t3);
}

template <bool INV> __device__ __forceinline__ void dft8(cd* a) {
  // radix-2 split into even/odd, then two radix-4
  cd e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cd o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  // twiddles w8^q on odd outputs: w8 = exp(-+ i pi/4)
  cd w1 = INV ? mk(kR * (o1.x - o1.y), kR * (o1.x + o1.y)) : mk(kR * (o1.x + o1.y), kR * (o1.y - o1.x));
  cd w2 = rot90<INV>(o2);
  cd w3 = INV ? mk(-kR * (o3.x + o3.y), kR * (o3.x - o3.y)) : mk(kR * (o3.y - o3.x), -kR * (o3.x + o3.y));
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, w1); a[5] = csub(e1, w1);
  a[2] = cadd(e2, w2); a[6] = csub(e2, w2);
  a[3] = cadd(e3, w3); a[7] = csub(e3, w3);
}

// 16-point DFT as 4 x 4: X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[4 n1 + n2] W4^(n1 k1)
template <bool INV> __device__ __forceinline__ void dft16(cd* a) {
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<INV>(a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12]);
  // a[n2 + 4 k1] *= W16^(n2 k1)
  a[5] = wmul<INV>(a[5], kC8, kS8);     // e = 1
  a[6] = wmul<INV>(a[6], kR, kR);       // e = 2
  a[7] = wmul<INV>(a[7], kS8, kC8);     // e = 3
  a[9] = wmul<INV>(a[9], kR, kR);       // e = 2
  a[10] = rot90<INV>(a[10]);            // e = 4
  a[11] = wmul<INV>(a[11], -kR, kR);    // e = 6
  a[13] = wmul<INV>(a[13], kS8, kC8);   // e = 3
  a[14] = wmul<INV>(a[14], -kR, kR);    // e = 6
  a[15] = wmul<INV>(a[15], -kC8, -kS8); // e = 9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<INV>(a[4 * k1], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
  // a[4 k1 + k2] holds X[k1 + 4 k2]: transpose
  cd b[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) b[k1 + 4 * k2] = a[4 * k1 + k2];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = b[k];
}

template <int R, bool INV> __device__ __forceinline__ void dftR(cd* a) {
  if constexpr (R == 16) dft16<INV>(a);
  else if constexpr (R == 8) dft8<INV>(a);
  else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
  else bfly2<INV>(a[0], a[1]);
}

template <bool INV> __device__ __forceinline__ cd twid(const cd* __restrict__ tw, int idx) {
  asm volatile("" : "+v"(idx));   // twiddle loads are not shared across transforms (frees their registers)
  cd w = tw[idx];
  return INV ? mk(w.x, -w.y) : w;
}

// LDS index with one padding element after every 2^CHNS_LDS_PAD: the
// Stockham writes at stride R (>= 128 B) would otherwise hit the same banks.
#ifndef CHNS_LDS_PAD
#define CHNS_LDS_PAD 3
#endif
__device__ __forceinline__ int lp(int i) { return i + (i >> CHNS_LDS_PAD); }
template <int LOG2N> struct LdsSize { static constexpr int value = (1 << LOG2N) + ((1 << LOG2N) >> CHNS_LDS_PAD); };

// cos and sin of 2 pi e/16, e < 16
__device__ __forceinline__ constexpr double cos16(int e) {
  constexpr double c[16] = {1.0, kC8, kR, kS8, 0.0, -kS8, -kR, -kC8, -1.0, -kC8, -kR, -kS8, 0.0, kS8, kR, kC8};
  return c[e & 15];
}
__device__ __forceinline__ constexpr double sin16(int e) { return cos16(e + 12); }
__device__ __forceinline__ cd csq(cd a) { return mk(a.x * a.x - a.y * a.y, 2.0 * (a.x * a.y)); }

// Entries of the per-workgroup LDS stage-twiddle table of an N-point
// transform with P points per thread: for each stage after the first, the
// base twiddles w = exp(-2 pi i k TWS/N) of its butterflies, k < NS for the
// radix-P stages, k < T for a smaller final radix (its butterflies u > 0 take
// k + u T = w(k) exp(-2 pi i u/P), a constant rotation).  Higher powers are
// formed by squaring.  N = 2048, P = 8: 8 + 64 + 256 = 328 entries (5.1 KiB).
template <int LOG2N, int P>
struct TwLds {
  static constexpr int N = 1 << LOG2N, T = N / P;
  static constexpr int entries(int NS) { return (N / NS) >= P ? NS : T; }
  static constexpr int off(int NS) {
    int o = 0;
    for (int s = P; s < NS; s *= P) o += entries(s);
    return o;
  }
  static constexpr int size = off(N);
  // fills the table from the global table tw[n] = exp(-2 pi i n/N); threads
  // tid < size of the workgroup take one entry each
  static __device__ __forceinline__ cd entry(const cd* __restrict__ tw, int e) {
    int NS = P, o = 0;
#pragma unroll
    for (int s = 0; s < LOG2N; ++s) {
      if (NS < N && e >= o + entries(NS)) { o += entries(NS); NS *= P; }
    }
    const int R = (N / NS) >= P ? P : (N / NS);
    return tw[(e - o) * (N / (NS * R))];
  }
};

// Twiddles of one stage for thread t: w, w^2, w^4 (, w^8) of each of its Q =
// P/R butterflies, w = exp(-+2 pi i k TWS/N), k = b mod NS.  They are loaded
// one stage ahead, before the LDS exchange that precedes their use, so the
// load latency hides behind the exchange.  LT = false: every power from the
// global table.  LT = true: the base power of butterfly 0 from the LDS table
// (TwLds), the others formed at use, so the transform issues no global load.
template <int LOG2N, int P, int NS, bool LT = false>
struct StageTw {
  static constexpr int N = 1 << LOG2N;
  static constexpr int R = (N / NS) >= P ? P : (N / NS);
  static constexpr int Q = P / R;
  static constexpr int NL = R >= 16 ? 4 : (R >= 8 ? 3 : (R >= 4 ? 2 : 1));
  cd w[LT ? 1 : Q][LT ? 1 : NL];
  template <bool INV>
  __device__ __forceinline__ void load(const cd* __restrict__ tw, int t) {
    if constexpr (NS > 1) {
      constexpr int T = N / P, TWS = N / (NS * R);
      if constexpr (LT) {
        w[0][0] = tw[TwLds<LOG2N, P>::off(NS) + (Q == 1 ? t % NS : t)];
      } else {
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          const int k = (t + u * T) % NS;
#pragma unroll
          for (int l = 0; l < NL; ++l) w[u][l] = twid<INV>(tw, (k << l) * TWS);
        }
      }
    }
  }
  // powers w^(2^l), l < NL, of butterfly u
  template <bool INV>
  __device__ __forceinline__ void powers(int u, cd* p) const {
    if constexpr (LT) {
      cd b = INV ? mk(w[0][0].x, -w[0][0].y) : w[0][0];
      if (u > 0) b = wmul<INV>(b, cos16(u * 16 / P), sin16(u * 16 / P));
      p[0] = b;
#pragma unroll
      for (int l = 1; l < NL; ++l) p[l] = csq(p[l - 1]);
    } else {
#pragma unroll
      for (int l = 0; l < NL; ++l) p[l] = w[u][l];
    }
  }
};

// One Stockham stage of radix R on the cyclic registers v[P] of thread t.
// NS = product of the radices of earlier stages.  LAST: outputs stay in v in
// cyclic order; otherwise outputs go to lds at their Stockham positions.
template <int LOG2N, int P, int R, int NS, bool INV, bool LAST, int S, bool PRUNED, class Tw>
__device__ __forceinline__ void stage(cd* v, cd* lds, const Tw& tws, int t) {
  constexpr int N = 1 << LOG2N;
  constexpr int T = N / P;
  constexpr int Q = P / R;            // butterflies per thread
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    const int b = t + u * T;
    const int k = b % NS;
    cd a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[u + m * Q];
    if constexpr (PRUNED && NS == 1 && R == P) {
      // dealiased input: points t + m*T with P/4 < m < 3P/4 lie in N/4 < n < 3N/4,
      // where every truncated spectrum is zero
#pragma unroll
      for (int m = P / 4 + 1; m < 3 * P / 4; ++m) a[m] = mk(0.0, 0.0);
    }
    if constexpr (NS > 1) {
      // w^m for m < R from w, w^2, w^4, w^8, applied as they are formed so
      // few powers are live at once
      cd pw[Tw::NL];
      tws.template powers<INV>(u, pw);
      const cd w1 = pw[0];
      a[1] = cmulc(a[1], w1);
      if constexpr (R >= 4) {
        const cd w2 = pw[1];
        const cd w3 = cmulc(w1, w2);
        a[2] = cmulc(a[2], w2);
        a[3] = cmulc(a[3], w3);
        if constexpr (R >= 8) {
          const cd w4 = pw[2];
          a[4] = cmulc(a[4], w4);
          a[5] = cmulc(a[5], cmulc(w1, w4));
          a[6] = cmulc(a[6], cmulc(w2, w4));
          a[7] = cmulc(a[7], cmulc(w3, w4));
          if constexpr (R == 16) {
            const cd w8 = pw[3];
            const cd w12 = cmulc(w4, w8);
            a[8] = cmulc(a[8], w8);
            a[9] = cmulc(a[9], cmulc(w1, w8));
            a[10] = cmulc(a[10], cmulc(w2, w8));
            a[11] = cmulc(a[11], cmulc(w3, w8));
            a[12] = cmulc(a[12], w12);
            a[13] = cmulc(a[13], cmulc(w1, w12));
            a[14] = cmulc(a[14], cmulc(w2, w12));
            a[15] = cmulc(a[15], cmulc(w3, w12));
          }
        }
      }
    }
    dftR<R, INV>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int q = 0; q < R; ++q) v[u + q * Q] = a[q];
    } else {
      const int base = (b / NS) * NS * R + k;
      // lp(base + q NS) = lp(base) + q NS + (q NS)/8: NS is 1 (then base = b R,
      // a multiple of 8) or a multiple of 8, so adding q NS carries its own
      // eighths; the writes then take constant offsets from one address
      static_assert(NS == 1 || NS % 8 == 0, "stage strides");
      static_assert(CHNS_LDS_PAD == 3, "offsets assume one pad per 8 entries");
      const int lb = lp(base);
#pragma unroll
      for (int q = 0; q < R; ++q) lds[(lb + q * NS + ((q * NS) >> 3)) * S] = a[q];
    }
  }
}

// Workgroup barrier for exchanges through LDS.  The fences order LDS accesses
// only (this wave's LDS accesses complete, lgkmcnt(0), before the barrier), so
// vector memory operations in flight -- the next operand of a pass, rows the
// DMA stager copies into its LDS buffers (issued as asm, which the compiler
// does not track) -- stay in flight across it; __syncthreads() would wait for
// every one of them (vmcnt(0)).  It makes no global memory and no DMA copy
// visible: a DMA buffer is read only after waiting for its copies (RowStager).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int LOG2N, int P, int S>
__device__ __forceinline__ void lds_read_cyclic(cd* v, const cd* lds, int t) {
  constexpr int T = (1 << LOG2N) / P;
  if constexpr (T % 8 == 0) {
    // lp(t + m T) = lp(t) + m (T + T/8): constant offsets from one address
    const int lt = lp(t);
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = lds[(lt + m * (T + T / 8)) * S];
  } else {
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = lds[lp(t + m * T) * S];
  }
}

template <int LOG2N, int P, bool INV, int S, int NS, bool PRUNED, bool LT>
__device__ __forceinline__ void fft_from(cd* v, cd* lds, const cd* __restrict__ tw, int t,
                                         const StageTw<LOG2N, P, NS, LT>& tws) {
  constexpr int N = 1 << LOG2N;
  constexpr int R = (N / NS) >= P ? P : (N / NS);
  constexpr bool LAST = NS * R == N;
  stage<LOG2N, P, R, NS, INV, LAST, S, PRUNED && NS == 1>(v, lds, tws, t);
  if constexpr (!LAST) {
    StageTw<LOG2N, P, NS * R, LT> nxt;
    nxt.template load<INV>(tw, t);   // in flight during the exchange
    lds_sync();
    lds_read_cyclic<LOG2N, P, S>(v, lds, t);
    lds_sync();
    fft_from<LOG2N, P, INV, S, NS * R, false, LT>(v, lds, tw, t, nxt);
  }
}

// Full transform.  v: cyclic input -> cyclic output (P points per thread).
// lds: this transform's slot in a workgroup array holding S transforms
// interleaved (entry e of transform c at lds_all[e*S + c]), so lanes of
// different transforms touch adjacent addresses instead of aliasing banks.
// Begins with a barrier, so back-to-back calls sharing lds are safe.
// PRUNED: the input is a dealiased (truncated) spectrum, zero for indices
// N/4 < n < 3N/4; the first stage skips those points.
// LT: tw is the workgroup's LDS stage-twiddle table (TwLds, filled before the
// first call) instead of the global table.
template <int LOG2N, int P, bool INV, int S, bool PRUNED = false, bool LT = false>
__device__ __forceinline__ void fft(cd* v, cd* lds, const cd* __restrict__ tw, int t) {
  static_assert(P == 8 || P == 16, "8 or 16 points per thread");
  static_assert((1 << LOG2N) >= P * 2 && LOG2N <= 12, "supported lengths");
  // t is made opaque per transform: otherwise the compiler hoists every stage's
  // LDS and twiddle index arithmetic (a function of t only) out of all the
  // transforms of a kernel and keeps it live throughout, costing ~100 VGPRs
  asm volatile("" : "+v"(t));
  lds_sync();
  const StageTw<LOG2N, P, 1, LT> none{};
  fft_from<LOG2N, P, INV, S, 1, PRUNED, LT>(v, lds, tw, t, none);
}


// ---------------------------------------------------------------------------
// 2048-point transform on one workgroup of 256 threads (4 waves), 8 points
// per thread, with a single workgroup barrier between its stages.
//
// 2048 = 4 x 512.  Inverse (decimation in frequency): thread t holds
// X[t + 256 j] (the cyclic spectral layout); with k = m + 512 r, the radix-4
// DFT over r of X[m + 512 r] for m in {t, t + 256} and the twiddle
// W^(-m q) are local to the thread; one exchange through LDS gives wave q
// the sequence z_q[m], and the wave transforms it with a 512-point Stockham
// FFT whose exchanges stay inside its own quarter of the LDS array (a wave's
// LDS accesses execute in order, so they need no workgroup barrier).  Lane l
// of wave q ends with x[4 (l + 64 j) + q].  The forward transform is the
// transpose (decimation in time): it takes x in that layout and returns
// X[t + 256 j].  So spectral data keeps the cyclic layout of fft() while
// real-space points sit at w4_pos(t, j) (see below).
constexpr int W4_Q = 512;   // LDS entries per wave quarter

// LDS index of entry i of a quarter: the low 3 bits XORed with bits 3-5, no
// padding.  Sixteen consecutive entries (the reads of every exchange) stay in
// sixteen distinct 16-byte bank groups, and so do eight entries at stride 8
// (the first stage's writes); the one-in-eight padding of lp() costs the
// consecutive reads a two-way conflict (measured on MI355X: 127 instead of
// 245 B/clk per CU for 16-byte reads; 16-byte writes run at ~76 B/clk either
// way).
__device__ __forceinline__ int lw(int i) { return i ^ ((i >> 3) & 7); }

__device__ __forceinline__ int w4_pos(int t, int j) { return 4 * (t & 63) + (t >> 6) + 256 * j; }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Twiddles of the 4-wave transform for thread t: eight entries of the
// 2048-point table, the inverse using their conjugates.
//   e = 0, 1: W^t, W^(2t)               (cross-wave radix-4 stage)
//   e = 2..4: W^(32k), W^(64k), W^(128k), k = l % 8   (wave stage NS = 8)
//   e = 5..7: W^(4l), W^(8l), W^(16l), l = t % 64     (wave stage NS = 64)
// HELD: loaded once (init) and kept in registers for every transform of the
// kernel, so no table load queues behind the kernel's row loads and stores
// (vector memory operations complete in issue order).  Otherwise each entry
// is loaded from the table where the transform first needs it.
template <bool HELD>
struct W4Tw {
  const cd* __restrict__ tw;
  cd w[8];
  static __device__ __forceinline__ int idx(int e, int t) {
    const int l = t & 63, k = l & 7;
    switch (e) {
      case 0: return t;
      case 1: return 2 * t;
      case 2: return 32 * k;
      case 3: return 64 * k;
      case 4: return 128 * k;
      case 5: return 4 * l;
      case 6: return 8 * l;
      default: return 16 * l;
    }
  }
  // WAIT: wait for the loads here, before the kernel issues its row copies
  // (DMA stager): later uses then carry no wait, which the compiler would
  // otherwise place at the first transform, behind every copy in flight.
  template <bool WAIT = false>
  __device__ __forceinline__ void init(const cd* __restrict__ table, int t) {
    tw = table;
    if constexpr (HELD) {
#pragma unroll
      for (int e = 0; e < 8; ++e) w[e] = tw[idx(e, t)];
      if constexpr (WAIT) {
#pragma unroll
        for (int e = 0; e < 8; ++e) asm volatile("" : "+v"(w[e].x), "+v"(w[e].y));
      }
    }
  }
  template <bool INV>
  __device__ __forceinline__ cd get(int e, int t) const {
    if constexpr (HELD) {
      // an opaque copy per use: products of twiddles (w^3, w^5, ...) are then
      // formed inside each transform instead of being kept live across all
      double x = w[e].x, y = w[e].y;
      asm volatile("" : "+v"(x), "+v"(y));
      return INV ? mk(x, -y) : mk(x, y);
    } else {
      return twid<INV>(tw, idx(e, t));
    }
  }
};

// 512-point Stockham transform by one wave: lane l = t % 64 holds points
// l + 64 m (cyclic in, cyclic out); lq is the wave's LDS quarter; twiddles are
// W_512^e = W^(4e) of the 2048-point table.
template <bool INV, bool PRUNED_ZERO_UPPER, bool HELD>
__device__ __forceinline__ void fft512_wave(cd* v, cd* lq, const W4Tw<HELD>& W, int t) {
  const int l = t & 63;
  // stage NS = 1 (no twiddles)
  {
    cd a[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) a[m] = v[m];
    dft8<INV>(a);
    const int base = l * 8;
    cd w1 = W.template get<INV>(2, t), w2 = W.template get<INV>(3, t), w4 = W.template get<INV>(4, t);
#pragma unroll
    for (int q = 0; q < 8; ++q) lq[lw(base + q)] = a[q];
    wave_sync();
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = lq[lw(l + 64 * m)];
    wave_sync();
    // stage NS = 8: k = l % 8, twiddle exponent k * 8 (W_512), table index 4 k 8
    a[0] = v[0];
    a[1] = cmulc(v[1], w1);
    const cd w3 = cmulc(w1, w2);
    a[2] = cmulc(v[2], w2);
    a[3] = cmulc(v[3], w3);
    a[4] = cmulc(v[4], w4);
    a[5] = cmulc(v[5], cmulc(w1, w4));
    a[6] = cmulc(v[6], cmulc(w2, w4));
    a[7] = cmulc(v[7], cmulc(w3, w4));
    dft8<INV>(a);
    const int k = l % 8;
    const int b2 = (l / 8) * 64 + k;
    w1 = W.template get<INV>(5, t);
    w2 = W.template get<INV>(6, t);
    w4 = W.template get<INV>(7, t);
#pragma unroll
    for (int q = 0; q < 8; ++q) lq[lw(b2 + q * 8)] = a[q];
    wave_sync();
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = lq[lw(l + 64 * m)];
    wave_sync();
    // stage NS = 64: k = l, twiddle exponent l (W_512), table index 4 l; last stage
    v[1] = cmulc(v[1], w1);
    const cd w3b = cmulc(w1, w2);
    v[2] = cmulc(v[2], w2);
    v[3] = cmulc(v[3], w3b);
    v[4] = cmulc(v[4], w4);
    v[5] = cmulc(v[5], cmulc(w1, w4));
    v[6] = cmulc(v[6], cmulc(w2, w4));
    v[7] = cmulc(v[7], cmulc(w3b, w4));
    dft8<INV>(v);
  }
}

// Inverse 2048-point transform (see above).  PRUNED: X[k] = 0 for 512 < k < 1536.
// SYNC = false: the caller has synchronised the workgroup since any other
// wave last accessed lds (the transform's first writes cross waves).
template <bool PRUNED, bool HELD, bool SYNC = true>
__device__ __forceinline__ void fft_w4_inv(cd* v, cd* lds, const W4Tw<HELD>& W, int t) {
  asm volatile("" : "+v"(t));
  if constexpr (SYNC) lds_sync();
  const cd c1 = mk(kR, kR);   // W_2048^(-256) = exp(+i pi/4)
  const cd w1t = W.template get<true>(0, t), w2t = W.template get<true>(1, t);
  if constexpr (PRUNED) {
    v[3] = mk(0.0, 0.0);
    v[4] = mk(0.0, 0.0);
    v[5] = mk(0.0, 0.0);
  }
#pragma unroll
  for (int jp = 0; jp < 2; ++jp) {
    const int m = t + 256 * jp;
    cd a0 = v[jp], a1 = v[jp + 2], a2 = v[jp + 4], a3 = v[jp + 6];
    dft4<true>(a0, a1, a2, a3);
    // twiddles W^(-m q): w(m) = w(t) c1^jp, w(2m) = w(2t) (+i)^jp
    cd wm1 = w1t, wm2 = w2t;
    if (jp == 1) {
      wm1 = cmulc(w1t, c1);
      wm2 = rot90<true>(w2t);
    }
    const cd wm3 = cmulc(wm1, wm2);
    lds[lw(m)] = a0;
    lds[W4_Q + lw(m)] = cmulc(a1, wm1);
    lds[2 * W4_Q + lw(m)] = cmulc(a2, wm2);
    lds[3 * W4_Q + lw(m)] = cmulc(a3, wm3);
  }
  lds_sync();
  const int q = t >> 6, l = t & 63;
  cd* lq = lds + q * W4_Q;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = lq[lw(l + 64 * j)];
  wave_sync();
  fft512_wave<true, false>(v, lq, W, t);
}

// Forward 2048-point transform: x in the w4_pos layout -> X[t + 256 j].
template <bool HELD>
__device__ __forceinline__ void fft_w4_fwd(cd* v, cd* lds, const W4Tw<HELD>& W, int t) {
  asm volatile("" : "+v"(t));
  lds_sync();
  const int q = t >> 6, l = t & 63;
  cd* lq = lds + q * W4_Q;
  fft512_wave<false, false>(v, lq, W, t);
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) lq[lw(l + 64 * j)] = v[j];
  const cd w1t = W.template get<false>(0, t), w2t = W.template get<false>(1, t);
  const cd c1 = mk(kR, -kR);   // W_2048^256 = exp(-i pi/4)
  lds_sync();
#pragma unroll
  for (int jp = 0; jp < 2; ++jp) {
    const int k = t + 256 * jp;
    cd wk1 = w1t, wk2 = w2t;
    if (jp == 1) {
      wk1 = cmulc(w1t, c1);
      wk2 = rot90<false>(w2t);
    }
    const cd wk3 = cmulc(wk1, wk2);
    cd b0 = lds[lw(k)];
    cd b1 = cmulc(lds[W4_Q + lw(k)], wk1);
    cd b2 = cmulc(lds[2 * W4_Q + lw(k)], wk2);
    cd b3 = cmulc(lds[3 * W4_Q + lw(k)], wk3);
    dft4<false>(b0, b1, b2, b3);
    v[jp] = b0;
    v[jp + 2] = b1;
    v[jp + 4] = b2;
    v[jp + 6] = b3;
  }
  lds_sync();   // callers (row_split) write the LDS array next
}

// Split of a paired real transform: Z = F(a + i b) -> A = F(a), B = F(b) at n
// from zn = Z[n] and zr = Z[N - n].
__device__ __forceinline__ void split_pair(cd zn, cd zr, cd* A, cd* B) {
  *A = mk((zn.x + zr.x) * 0.5, (zn.y - zr.y) * 0.5);
  *B = mk((zn.y + zr.y) * 0.5, (zr.x - zn.x) * 0.5);
}

// Forward 2048-point transform of the w4_pos layout fused with the split of a
// paired real transform: hands A = F(a), B = F(b) for every n <= N/2 to
// op(n, A, B), nothing in v.  The final radix-4 stage gives thread t the
// positions k = t and k' = 512 - t (thread 0: 0 and 256): the mirror
// N - (k + 512 q) of each output it forms is then k' + 512 (3 - q) (thread 0:
// its own), so no exchange through LDS is needed for the split.  Twiddles of
// k': W^(512 - t) = -i conj(W^t), W^(1024 - 2t) = -conj(W^(2t)).
// SYNC = false: no other wave has accessed this wave's quarter of lds since
// the caller's last workgroup barrier (the first stages stay in the quarter).
template <bool HELD, bool SYNC = true, class Op>
__device__ __forceinline__ void fft_w4_fwd_split(cd* v, cd* lds, const W4Tw<HELD>& W, int t, Op op) {
  asm volatile("" : "+v"(t));
  if constexpr (SYNC) lds_sync();
  const int q = t >> 6, l = t & 63;
  cd* lq = lds + q * W4_Q;
  fft512_wave<false, false>(v, lq, W, t);
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; ++j) lq[lw(l + 64 * j)] = v[j];
  const cd w1t = W.template get<false>(0, t), w2t = W.template get<false>(1, t);
  lds_sync();
  const bool t0 = t == 0;
  const int k2 = t0 ? 256 : 512 - t;
  cd x[4], y[4];
  {
    const cd w3 = cmulc(w1t, w2t);
    x[0] = lds[lw(t)];
    x[1] = cmulc(lds[W4_Q + lw(t)], w1t);
    x[2] = cmulc(lds[2 * W4_Q + lw(t)], w2t);
    x[3] = cmulc(lds[3 * W4_Q + lw(t)], w3);
    dft4<false>(x[0], x[1], x[2], x[3]);   // X[t + 512 q]
  }
  {
    const cd u1 = t0 ? mk(kR, -kR) : mk(-w1t.y, -w1t.x);     // W^256 = exp(-i pi/4)
    const cd u2 = t0 ? mk(0.0, -1.0) : mk(-w2t.x, w2t.y);    // W^512 = -i
    const cd u3 = cmulc(u1, u2);
    y[0] = lds[lw(k2)];
    y[1] = cmulc(lds[W4_Q + lw(k2)], u1);
    y[2] = cmulc(lds[2 * W4_Q + lw(k2)], u2);
    y[3] = cmulc(lds[3 * W4_Q + lw(k2)], u3);
    dft4<false>(y[0], y[1], y[2], y[3]);   // X[k2 + 512 q]
  }
  cd A, B;
  split_pair(x[0], t0 ? x[0] : y[3], &A, &B);
  op(t, A, B);
  split_pair(x[1], t0 ? x[3] : y[2], &A, &B);
  op(t + 512, A, B);
  split_pair(y[0], t0 ? y[3] : x[3], &A, &B);
  op(k2, A, B);
  split_pair(y[1], t0 ? y[2] : x[2], &A, &B);
  op(k2 + 512, A, B);
  if (t0) {
    split_pair(x[2], x[2], &A, &B);
    op(1024, A, B);
  }
}

// LDS entries of a row kernel's transform array: the four swizzled wave
// quarters of the 2048-point transform, R padded rows otherwise.
template <int LOG2N, int R> struct RowLds {
  static constexpr int value = (LOG2N == 11 && R == 1) ? 4 * W4_Q : R * LdsSize<LOG2N>::value;
};

// Row transform of a kernel with R rows per workgroup (rows interleaved
// across lanes, 8 points per thread): the 4-wave transform at N = 2048 with
// one row per workgroup, fft() otherwise.  rpos gives the real-space index of
// register j of thread t.
template <int LOG2N, bool INV, int R, bool PRUNED = false>
__device__ __forceinline__ void rfft(cd* v, cd* lds, const cd* __restrict__ tw, int t) {
  if constexpr (LOG2N == 11 && R == 1) {
    W4Tw<false> W;
    W.init(tw, t);
    if constexpr (INV) fft_w4_inv<PRUNED>(v, lds, W, t);
    else fft_w4_fwd(v, lds, W, t);
  } else {
    fft<LOG2N, 8, INV, R, PRUNED>(v, lds, tw, t);
  }
}

// Twiddles of a row kernel whose transforms share them: held in registers for
// the 4-wave transform (init loads them; call it before the kernel's row
// loads), the table otherwise.
template <int LOG2N, int R>
struct RowTw {
  static constexpr bool W4 = LOG2N == 11 && R == 1;
  W4Tw<W4> w;
  template <bool WAIT = false>
  __device__ __forceinline__ void init(const cd* __restrict__ tw, int t) { w.template init<WAIT>(tw, t); }
};
// SYNC = false skips the opening barrier of the 4-wave inverse transform
// (see fft_w4_inv); the other transforms always synchronise.
template <int LOG2N, bool INV, int R, bool PRUNED = false, bool SYNC = true>
__device__ __forceinline__ void rfft(cd* v, cd* lds, const RowTw<LOG2N, R>& T, int t) {
  if constexpr (RowTw<LOG2N, R>::W4) {
    if constexpr (INV) fft_w4_inv<PRUNED, true, SYNC>(v, lds, T.w, t);
    else fft_w4_fwd(v, lds, T.w, t);
  } else {
    fft<LOG2N, 8, INV, R, PRUNED>(v, lds, T.w.tw, t);
  }
}
template <int LOG2N, int R>
__device__ __forceinline__ int rpos(int t, int j) {
  if constexpr (LOG2N == 11 && R == 1) return w4_pos(t, j);
  else return t + j * ((1 << LOG2N) / 8);
}

}  // namespace chns

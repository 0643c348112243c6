// FP64 Stockham FFT engine for gfx950, staged through LDS.
//
// One transform of length N = 2^LOG2N is carried by T = N/8 threads; thread t
// holds the eight points t + m*T (m = 0..7) in registers ("cyclic"
// distribution) both before and after the transform, so a c2r can feed a
// pointwise stage and an r2c without any extra data movement.
//
// Stages are radix 8 (registers) with a final radix 2 or 4 when LOG2N is not a
// multiple of 3.  Between stages the values go through LDS: every thread
// writes its butterfly outputs to their Stockham positions, the workgroup
// synchronises, and every thread reads its cyclic set back.  The
// synchronisations are workgroup-wide, so all threads of a workgroup must call
// the same transforms in the same order (inputs of idle lanes are zero).
//
// Twiddles come from a table tw[n] = exp(-2*pi*i*n/N), n < N, built on the
// host in extended precision; the inverse uses the conjugates.  No
// normalisation is applied (the 1/(Nx*Ny) of the c2r is applied by callers).
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
  constexpr double R = 0.70710678118654752440084436210485;
  // radix-2 split into even/odd, then two radix-4
  cd e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cd o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  // twiddles w8^q on odd outputs: w8 = exp(-+ i pi/4)
  cd w1 = INV ? mk(R * (o1.x - o1.y), R * (o1.x + o1.y)) : mk(R * (o1.x + o1.y), R * (o1.y - o1.x));
  cd w2 = rot90<INV>(o2);
  cd w3 = INV ? mk(-R * (o3.x + o3.y), R * (o3.x - o3.y)) : mk(R * (o3.y - o3.x), -R * (o3.x + o3.y));
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, w1); a[5] = csub(e1, w1);
  a[2] = cadd(e2, w2); a[6] = csub(e2, w2);
  a[3] = cadd(e3, w3); a[7] = csub(e3, w3);
}

template <int R, bool INV> __device__ __forceinline__ void dftR(cd* a) {
  if constexpr (R == 8) dft8<INV>(a);
  else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
  else bfly2<INV>(a[0], a[1]);
}

template <bool INV> __device__ __forceinline__ cd twid(const cd* __restrict__ tw, int idx) {
  cd w = tw[idx];
  return INV ? mk(w.x, -w.y) : w;
}

// LDS index with one padding element after every 8: the radix-8 Stockham
// writes at stride 8 (128 B) would otherwise hit the same banks.
__device__ __forceinline__ int lp(int i) { return i + (i >> 3); }
template <int LOG2N> struct LdsSize { static constexpr int value = (1 << LOG2N) + (1 << LOG2N) / 8; };

// One Stockham stage of radix R on the cyclic registers v[8] of thread t.
// Ns = product of the radices of earlier stages.  LAST: outputs stay in v in
// cyclic order; otherwise outputs go to lds at their Stockham positions.
template <int LOG2N, int R, int NS, bool INV, bool LAST, int S>
__device__ __forceinline__ void stage(cd* v, cd* lds, const cd* __restrict__ tw, int t) {
  constexpr int N = 1 << LOG2N;
  constexpr int T = N / 8;
  constexpr int P = 8 / R;            // butterflies per thread
  constexpr int TWS = N / (NS * R);   // twiddle index stride
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const int b = t + u * T;
    const int k = b % NS;
    cd a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[u + m * P];
    if constexpr (NS > 1) {
      // w^m for m < R from three table loads (w, w^2, w^4)
      cd w[8];
      w[1] = twid<INV>(tw, k * TWS);
      if constexpr (R >= 4) {
        w[2] = twid<INV>(tw, 2 * k * TWS);
        w[3] = cmulc(w[1], w[2]);
      }
      if constexpr (R == 8) {
        w[4] = twid<INV>(tw, 4 * k * TWS);
        w[5] = cmulc(w[1], w[4]);
        w[6] = cmulc(w[2], w[4]);
        w[7] = cmulc(w[3], w[4]);
      }
#pragma unroll
      for (int m = 1; m < R; ++m) a[m] = cmulc(a[m], w[m]);
    }
    dftR<R, INV>(a);
    if constexpr (LAST) {
#pragma unroll
      for (int q = 0; q < R; ++q) v[u + q * P] = a[q];
    } else {
      const int base = (b / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) lds[lp(base + q * NS) * S] = a[q];
    }
  }
}

__device__ __forceinline__ void lds_sync() { __syncthreads(); }

template <int LOG2N, int S>
__device__ __forceinline__ void lds_read_cyclic(cd* v, const cd* lds, int t) {
  constexpr int T = (1 << LOG2N) / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = lds[lp(t + m * T) * S];
}

// Full transform.  v: cyclic input -> cyclic output.  lds: this transform's
// slot in a workgroup array holding S transforms interleaved (entry e of
// transform c at lds_all[e*S + c]), so lanes of different transforms touch
// adjacent addresses instead of aliasing banks.  Starts and ends with a workgroup barrier-free state:
// the caller must synchronise before lds is reused by a different transform
// (fft() itself begins with a barrier so back-to-back calls are safe).
template <int LOG2N, bool INV, int S>
__device__ __forceinline__ void fft(cd* v, cd* lds, const cd* __restrict__ tw, int t) {
  static_assert(LOG2N >= 4 && LOG2N <= 12, "supported lengths 16..4096");
  constexpr int NR8 = LOG2N / 3;
  constexpr int REM = LOG2N % 3;
  // stage 1 .. NR8 radix 8, then radix 2^REM
  lds_sync();
  if constexpr (REM == 0) {
    // NR8 >= 2 here
    stage<LOG2N, 8, 1, INV, false, S>(v, lds, tw, t);
    lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
    if constexpr (NR8 == 2) {
      stage<LOG2N, 8, 8, INV, true, S>(v, lds, tw, t);
    } else if constexpr (NR8 == 3) {
      stage<LOG2N, 8, 8, INV, false, S>(v, lds, tw, t);
      lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
      stage<LOG2N, 8, 64, INV, true, S>(v, lds, tw, t);
    } else {
      stage<LOG2N, 8, 8, INV, false, S>(v, lds, tw, t);
      lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
      stage<LOG2N, 8, 64, INV, false, S>(v, lds, tw, t);
      lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
      stage<LOG2N, 8, 512, INV, true, S>(v, lds, tw, t);
    }
  } else {
    constexpr int RL = 1 << REM;
    stage<LOG2N, 8, 1, INV, false, S>(v, lds, tw, t);
    lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
    if constexpr (NR8 >= 2) {
      stage<LOG2N, 8, 8, INV, false, S>(v, lds, tw, t);
      lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
    }
    if constexpr (NR8 >= 3) {
      stage<LOG2N, 8, 64, INV, false, S>(v, lds, tw, t);
      lds_sync(); lds_read_cyclic<LOG2N, S>(v, lds, t); lds_sync();
    }
    constexpr int NSL = (NR8 == 1) ? 8 : (NR8 == 2 ? 64 : 512);
    stage<LOG2N, RL, NSL, INV, true, S>(v, lds, tw, t);
  }
}

}  // namespace chns

// Register-resident Stockham FFT for power-of-two lengths 8..4096.
//
// T = N/8 threads cooperate on one transform; each thread holds 8 elements in
// registers.  Stages are radix 8 (plus one radix-2 or radix-4 stage when
// log2 N is not a multiple of 3).  Stage s (radix R, L = product of earlier
// radices) computes, for butterfly b (k = b mod L), the R inputs
// x[b + r*N/R] times the twiddles w_{LR}^{r k}, an R-point DFT, and writes
// y[(b - k) R + k + q L].  Thread j owns butterflies b = j + i*T.
// Input and output use the same "strided" layout: register m of thread j is
// element j + m*T, so global loads and stores are coalesced across threads.
// Between stages the data goes through LDS with one padding slot per 16
// elements against bank conflicts.  When a transform fits in one wavefront
// (N <= 512) the exchange needs no barrier: a wave's LDS operations execute
// in program order, so each stage reads and then overwrites one private
// buffer.  Larger transforms span several waves and use ping-pong buffers
// with one __syncthreads() per exchange.
#pragma once

#include "odt_device.h"

namespace odt {

__host__ __device__ constexpr int ilog2c(int n) {
  return n <= 1 ? 0 : 1 + ilog2c(n >> 1);
}

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_padded(int n) { return n + (n >> 4) + 1; }

// e^{dir * i * pi/4} etc. for the radix-8 butterfly
template <typename F, int DIR>
__device__ __forceinline__ cpx<F> mul_mi(cpx<F> a) {  // a * (dir*i)
  return DIR > 0 ? mk<F>(-a.im, a.re) : mk<F>(a.im, -a.re);
}

template <typename F, int DIR>
__device__ __forceinline__ void dft2(cpx<F> &a0, cpx<F> &a1) {
  const cpx<F> t = a0 - a1;
  a0 = a0 + a1;
  a1 = t;
}

template <typename F, int DIR>
__device__ __forceinline__ void dft4(cpx<F> &x0, cpx<F> &x1, cpx<F> &x2,
                                     cpx<F> &x3) {
  const cpx<F> t0 = x0 + x2, t1 = x0 - x2;
  const cpx<F> t2 = x1 + x3, t3 = mul_mi<F, DIR>(x1 - x3);
  x0 = t0 + t2;
  x2 = t0 - t2;
  x1 = t1 + t3;
  x3 = t1 - t3;
}

template <typename F, int DIR>
__device__ __forceinline__ void dft8(cpx<F> *a) {
  // even/odd split: E = DFT4(a0,a2,a4,a6), O = DFT4(a1,a3,a5,a7)
  cpx<F> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cpx<F> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<F, DIR>(e0, e1, e2, e3);
  dft4<F, DIR>(o0, o1, o2, o3);
  const F h = (F)0.70710678118654752440;
  // w8^1 = (1 + dir*i)/sqrt2, w8^2 = dir*i, w8^3 = (-1 + dir*i)/sqrt2
  const cpx<F> w1o1 = DIR > 0 ? mk<F>(h * (o1.re - o1.im), h * (o1.re + o1.im))
                              : mk<F>(h * (o1.re + o1.im), h * (o1.im - o1.re));
  const cpx<F> w2o2 = mul_mi<F, DIR>(o2);
  const cpx<F> w3o3 = DIR > 0
                          ? mk<F>(-h * (o3.re + o3.im), h * (o3.re - o3.im))
                          : mk<F>(h * (o3.im - o3.re), -h * (o3.re + o3.im));
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + w1o1;
  a[5] = e1 - w1o1;
  a[2] = e2 + w2o2;
  a[6] = e2 - w2o2;
  a[3] = e3 + w3o3;
  a[7] = e3 - w3o3;
}

// dft8 with a[2..5] = 0 (first stage of a transform whose nonzero inputs lie
// in the signed-frequency band (-N/4, N/4)); a[2..5] are not read.
template <typename F, int DIR>
__device__ __forceinline__ void dft8_band(cpx<F> *a) {
  const cpx<F> m6 = mul_mi<F, DIR>(a[6]), m7 = mul_mi<F, DIR>(a[7]);
  const cpx<F> e0 = a[0] + a[6], e2 = a[0] - a[6], e1 = a[0] - m6,
               e3 = a[0] + m6;
  const cpx<F> o0 = a[1] + a[7], o2 = a[1] - a[7], o1 = a[1] - m7,
               o3 = a[1] + m7;
  const F h = (F)0.70710678118654752440;
  const cpx<F> w1o1 = DIR > 0 ? mk<F>(h * (o1.re - o1.im), h * (o1.re + o1.im))
                              : mk<F>(h * (o1.re + o1.im), h * (o1.im - o1.re));
  const cpx<F> w2o2 = mul_mi<F, DIR>(o2);
  const cpx<F> w3o3 = DIR > 0
                          ? mk<F>(-h * (o3.re + o3.im), h * (o3.re - o3.im))
                          : mk<F>(h * (o3.im - o3.re), -h * (o3.re + o3.im));
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + w1o1;
  a[5] = e1 - w1o1;
  a[2] = e2 + w2o2;
  a[6] = e2 - w2o2;
  a[3] = e3 + w3o3;
  a[7] = e3 - w3o3;
}

template <typename F, int R, int DIR>
__device__ __forceinline__ void dft_r(cpx<F> *a) {
  if (R == 8) {
    dft8<F, DIR>(a);
  } else if (R == 4) {
    dft4<F, DIR>(a[0], a[1], a[2], a[3]);
  } else {
    dft2<F, DIR>(a[0], a[1]);
  }
}

template <typename F, int N, int DIR>
struct Fft8 {
  static constexpr int T = N / 8;
  static constexpr int LOG = ilog2c(N);
  static constexpr int N8 = LOG / 3;
  static constexpr int REM = LOG % 3;
  static constexpr int NS = N8 + (REM ? 1 : 0);
  static constexpr int PADN = lds_padded(N);
  static __host__ __device__ constexpr int radix(int s) {
    return s < N8 ? 8 : (REM == 1 ? 2 : 4);
  }

  // Twiddles of one thread, all stages: stage s >= 1 (radix R, L earlier
  // points) needs w_{LR}^{r k} for r = 1..R-1 and each of its 8/R
  // butterflies (k = b mod L, b = j + i*T).  They depend on j only, so a
  // thread that runs many transforms loads them once (Twid, load_twiddles).
  static __host__ __device__ constexpr int tw_count(int s) {
    return s == 0 ? 0 : (8 / radix(s)) * (radix(s) - 1);
  }
  static __host__ __device__ constexpr int tw_off(int s) {
    return s <= 1 ? 0 : tw_off(s - 1) + tw_count(s - 1);
  }
  static constexpr int NTW = tw_off(NS);
  struct Twid {
    cpx<F> w[NTW > 0 ? NTW : 1];
  };

  template <int S, int L>
  __device__ static __forceinline__ void load_tw(Twid &t,
                                                 const cpx<F> *__restrict__ tw,
                                                 int j) {
    constexpr int R = radix(S);
    constexpr int NB = 8 / R;
    if constexpr (S > 0) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int k = (j + i * T) & (L - 1);
#pragma unroll
        for (int r = 1; r < R; ++r)
          t.w[tw_off(S) + i * (R - 1) + r - 1] =
              conj_if(tw[r * k * (N / (L * R))], DIR < 0);
      }
    }
    if constexpr (S + 1 < NS) load_tw<S + 1, L * R>(t, tw, j);
  }
  __device__ static __forceinline__ void load_twiddles(Twid &t,
                                                       const cpx<F> *tw,
                                                       int j) {
    load_tw<0, 1>(t, tw, j);
  }

  template <int S, int L, bool BAND, bool REG>
  __device__ static __forceinline__ void stage(cpx<F> *v, cpx<F> *buf0,
                                               cpx<F> *buf1,
                                               const cpx<F> *__restrict__ tw,
                                               const Twid *rt, int j) {
    constexpr int R = radix(S);
    constexpr int NB = 8 / R;
    constexpr bool FIRST = S == 0;
    constexpr bool LAST = S == NS - 1;
    constexpr bool WAVE = T <= 64;
    cpx<F> *src = WAVE ? buf0 : ((S & 1) ? buf0 : buf1);  // from stage S-1
    cpx<F> *dst = WAVE ? buf0 : ((S & 1) ? buf1 : buf0);
    cpx<F> a[8];
    constexpr bool PRUNED = BAND && FIRST && R == 8;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!PRUNED || r < 2 || r > 5)
          a[i * R + r] = FIRST ? v[i + r * NB] : src[lds_pad(j + i * T + r * (N / R))];
    // wave mode: no compiler motion of this stage's LDS writes above its reads
    if (WAVE && !FIRST) __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = j + i * T;
      const int k = b & (L - 1);
      if (L > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cpx<F> w = REG ? rt->w[tw_off(S) + i * (R - 1) + r - 1]
                               : conj_if(tw[r * k * (N / (L * R))], DIR < 0);
          a[i * R + r] = a[i * R + r] * w;
        }
      }
      if constexpr (PRUNED)
        dft8_band<F, DIR>(a + i * R);
      else
        dft_r<F, R, DIR>(a + i * R);
    }
    if (LAST) {
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int q = 0; q < R; ++q) v[i + q * NB] = a[i * R + q];
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int b = j + i * T;
        const int k = b & (L - 1);
#pragma unroll
        for (int q = 0; q < R; ++q)
          dst[lds_pad((b - k) * R + k + q * L)] = a[i * R + q];
      }
      if (WAVE)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      else
        __syncthreads();
    }
  }

  template <int S, int L, bool BAND, bool REG>
  __device__ static __forceinline__ void stages(cpx<F> *v, cpx<F> *b0,
                                                cpx<F> *b1, const cpx<F> *tw,
                                                const Twid *rt, int j) {
    stage<S, L, BAND, REG>(v, b0, b1, tw, rt, j);
    if constexpr (S + 1 < NS)
      stages<S + 1, L * radix(S), BAND, REG>(v, b0, b1, tw, rt, j);
  }

  // Buffers one transform needs (PADN elements each).
  static constexpr int NBUF = T <= 64 ? 1 : 2;

  // v[m] = x[j + m*T] on entry, X[j + m*T] on exit.  buf0 (and buf1 when
  // NBUF == 2): this transform's LDS scratch.  For NBUF == 2 all threads of
  // the block must call it (it contains barriers).
  // BAND: the nonzero inputs lie in the signed-frequency band (-N/4, N/4),
  // i.e. v[2..5] are zero (and are not read).
  template <bool BAND = false>
  __device__ static __forceinline__ void run(cpx<F> *v, cpx<F> *buf0,
                                             cpx<F> *buf1, const cpx<F> *tw,
                                             int j) {
    // the previous call's last stage may still be reading the buffers
    if (NBUF == 2 && NS > 1) __syncthreads();
    stages<0, 1, BAND, false>(v, buf0, buf1, tw, nullptr, j);
  }
  // Same with this thread's twiddles already in registers.
  template <bool BAND = false>
  __device__ static __forceinline__ void run(cpx<F> *v, cpx<F> *buf0,
                                             cpx<F> *buf1, const Twid &t,
                                             int j) {
    if (NBUF == 2 && NS > 1) __syncthreads();
    stages<0, 1, BAND, true>(v, buf0, buf1, nullptr, &t, j);
  }
};

}  // namespace odt

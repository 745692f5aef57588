// Device building blocks for the gfx950 filtered-backpropagation kernels:
// complex arithmetic, workgroup-cooperative LDS transforms (Stockham radix-2
// for powers of two, Bluestein's chirp convolution otherwise), B-spline tap
// weights and the mirror index rule of scipy.ndimage (the resampler the
// reference calls in odtbrain/_alg3d_bpp.py:73-81 and _alg3d_bppt.py:931-944).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace odt {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cpx<R> mk(R a, R b) {
  cpx<R> c;
  c.re = a;
  c.im = b;
  return c;
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return mk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator*(R s, cpx<R> a) {
  return mk<R>(s * a.re, s * a.im);
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> conj_if(cpx<R> a, bool c) {
  return c ? mk<R>(a.re, -a.im) : a;
}

// ---------------------------------------------------------------------------
// LDS transforms.  A workgroup transforms `nb` vectors of length n stored
// contiguously in LDS buffer x (x and the scratch y each hold nb * lds_len(n)
// entries).  For powers of two tw holds exp(+2 pi i m / n), m = 0..n-1
// (other lengths: see lds_bluestein); dir < 0 uses the conjugate (forward).
// Returns the buffer holding the result.  Both end with __syncthreads().
// ---------------------------------------------------------------------------

// Stockham autosort radix-2: pass s combines sub-transforms of length
// Ls = 2^s; butterfly j (k = j mod Ls) reads x[j], x[j + n/2] and writes
// y[2(j-k) + k] and y[2(j-k) + k + Ls] with twiddle exp(i*dir*pi*k/Ls).
template <typename R>
__device__ cpx<R> *lds_fft_pow2(cpx<R> *x, cpx<R> *y, int n, int logn, int nb,
                                const cpx<R> *__restrict__ tw, int dir) {
  const int half = n >> 1;
  const int total = half * nb;
  for (int s = 0; s < logn; ++s) {
    const int ls = 1 << s;
    const int tstride = n >> (s + 1);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int b = idx >> (logn - 1);
      const int j = idx & (half - 1);
      const int k = j & (ls - 1);
      const cpx<R> a = x[b * n + j];
      const cpx<R> c = x[b * n + j + half];
      const cpx<R> w = conj_if(tw[k * tstride], dir < 0);
      const cpx<R> wc = w * c;
      const int o = b * n + ((j - k) << 1) + k;
      y[o] = a + wc;
      y[o + ls] = a - wc;
    }
    __syncthreads();
    cpx<R> *t = x;
    x = y;
    y = t;
  }
  return x;
}

// Lengths that are not powers of two (the no-padding case, where the
// reference's FFTW plans take any size) use Bluestein's algorithm up to
// kBluesteinMax: with c_j = exp(s i pi j^2 / n),
//   X_k = sum_j x_j exp(s 2 pi i jk / n) = c_k sum_j (x_j c_j) conj(c_{k-j}),
// a linear convolution evaluated circularly with power-of-two transforms of
// length m >= 2n - 1.  Longer ones take the direct DFT.  The twiddle table
// of such a length holds 2m + n entries: exp(2 pi i k / m) (k < m), the
// chirp w_j = exp(i pi j^2 / n) (j < n), and H = FFT_m(h) / m of the
// kernel h_d = conj(w_|d|) placed at d and m - d (the s = +1 kernel; the
// s = -1 one is its conjugate, and H is even, so its transform is conj(H)).
constexpr int kBluesteinMax = 2048;

__host__ __device__ __forceinline__ bool is_pow2(int n) {
  return n > 0 && (n & (n - 1)) == 0;
}
__host__ __device__ __forceinline__ int bluestein_len(int n) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  return m;
}
__host__ __device__ __forceinline__ bool uses_bluestein(int n) {
  return !is_pow2(n) && n <= kBluesteinMax;
}
// LDS entries per vector that the transform of length n works in
__host__ __device__ __forceinline__ int lds_len(int n) {
  return uses_bluestein(n) ? bluestein_len(n) : n;
}
// entries of the twiddle table of length n
__host__ __device__ __forceinline__ int dft_table_len(int n) {
  return uses_bluestein(n) ? 2 * bluestein_len(n) + n : n;
}

// Bluestein transform of nb vectors of length n (stride n in x); x and y
// each hold nb * bluestein_len(n) entries.  Returns the buffer with the
// result at stride n.
template <typename R>
__device__ cpx<R> *lds_bluestein(cpx<R> *x, cpx<R> *y, int n, int nb,
                                 const cpx<R> *__restrict__ tw, int dir) {
  const int m = bluestein_len(n);
  const int logm = 31 - __clz(m);
  const cpx<R> *chirp = tw + m, *H = tw + m + n;
  const bool fwd = dir < 0;
  for (int idx = threadIdx.x; idx < nb * m; idx += blockDim.x) {
    const int b = idx >> logm, j = idx & (m - 1);
    y[idx] = j < n ? x[b * n + j] * conj_if(chirp[j], fwd) : mk<R>(0, 0);
  }
  __syncthreads();
  cpx<R> *z = lds_fft_pow2<R>(y, x, m, logm, nb, tw, -1);
  for (int idx = threadIdx.x; idx < nb * m; idx += blockDim.x)
    z[idx] = z[idx] * conj_if(H[idx & (m - 1)], fwd);
  __syncthreads();
  z = lds_fft_pow2<R>(z, z == y ? x : y, m, logm, nb, tw, +1);
  cpx<R> *o = z == y ? x : y;
  for (int idx = threadIdx.x; idx < nb * n; idx += blockDim.x) {
    const int b = idx / n, k = idx - b * n;
    o[idx] = z[b * m + k] * conj_if(chirp[k], fwd);
  }
  __syncthreads();
  return o;
}

// Direct DFT for the lengths beyond kBluesteinMax that are not powers of two.
template <typename R>
__device__ cpx<R> *lds_dft(cpx<R> *x, cpx<R> *y, int n, int nb,
                           const cpx<R> *__restrict__ tw, int dir) {
  for (int idx = threadIdx.x; idx < n * nb; idx += blockDim.x) {
    const int b = idx / n;
    const int m = idx - b * n;
    cpx<R> acc = mk<R>(0, 0);
    int p = 0;
    for (int j = 0; j < n; ++j) {
      acc = acc + x[b * n + j] * conj_if(tw[p], dir < 0);
      p += m;
      if (p >= n) p -= n;
    }
    y[idx] = acc;
  }
  __syncthreads();
  return y;
}

template <typename R>
__device__ __forceinline__ cpx<R> *lds_transform(cpx<R> *x, cpx<R> *y, int n,
                                                 int logn, int nb,
                                                 const cpx<R> *tw, int dir) {
  if (logn >= 0) return lds_fft_pow2<R>(x, y, n, logn, nb, tw, dir);
  if (uses_bluestein(n)) return lds_bluestein<R>(x, y, n, nb, tw, dir);
  return lds_dft<R>(x, y, n, nb, tw, dir);
}

// Index j taken modulo n into [0, n).
__device__ __forceinline__ int wrap(int j, int n) {
  j %= n;
  return j < 0 ? j + n : j;
}

// ---------------------------------------------------------------------------
// scipy.ndimage resampling rules (restated in oracle/bspline.py).
// ---------------------------------------------------------------------------

// Whole-sample-symmetric extension of a tap index onto [0, n-1].
__device__ __forceinline__ int mirror_idx(int idx, int n) {
  if (n <= 1) return 0;
  const int s2 = 2 * n - 2;
  if (idx < 0) {
    idx = s2 * (-idx / s2) + idx;
    idx = idx <= 1 - n ? idx + s2 : -idx;
  } else if (idx >= n) {
    idx -= s2 * (idx / s2);
    if (idx >= n) idx = s2 - idx;
  }
  return idx;
}

// First tap and weights of the centred B-spline of degree ORDER at
// coordinate c (float64).  Taps are start .. start+ORDER.
template <int ORDER>
__device__ __forceinline__ int spline_start(double c) {
  return (ORDER & 1) ? (int)floor(c) - ORDER / 2
                     : (int)floor(c + 0.5) - ORDER / 2;
}

template <int ORDER, typename W>
__device__ __forceinline__ void spline_weights(double c, W *w) {
  const double t = (ORDER & 1) ? c - floor(c) : c - floor(c + 0.5);
  if (ORDER == 0) {
    w[0] = W(1);
  } else if (ORDER == 1) {
    w[0] = W(1.0 - t);
    w[1] = W(t);
  } else if (ORDER == 2) {
    // t in [-0.5, 0.5): taps at distance 1+t, t, 1-t
    const double a = 0.5 - t;
    w[0] = W(0.5 * a * a);
    w[1] = W(0.75 - t * t);
    w[2] = W(1.0 - 0.5 * a * a - (0.75 - t * t));
  } else if (ORDER == 3) {
    // t in [0, 1): taps at distance 1+t, t, 1-t, 2-t
    const double z = 1.0 - t;
    const double w1 = (t * t * (t - 2.0) * 3.0 + 4.0) / 6.0;
    const double w2 = (z * z * (z - 2.0) * 3.0 + 4.0) / 6.0;
    const double w0 = z * z * z / 6.0;
    w[0] = W(w0);
    w[1] = W(w1);
    w[2] = W(w2);
    w[3] = W(1.0 - w0 - w1 - w2);
  } else if (ORDER == 4) {
    // distances d = 2+t, 1+t, t, 1-t, 2-t with t in [-0.5, 0.5)
    double s = 0.0;
    for (int j = 0; j < 4; ++j) {
      const double d = fabs(t + 2.0 - j);
      double v;
      if (d < 0.5) {
        const double d2 = d * d;
        v = d2 * (d2 * 0.25 - 0.625) + 115.0 / 192.0;
      } else if (d < 1.5) {
        v = (55.0 + d * (20.0 + d * (-120.0 + d * (80.0 - 16.0 * d)))) /
            96.0;
      } else {
        const double e = 2.5 - d;
        v = e * e * e * e / 24.0;
      }
      w[j] = W(v);
      s += v;
    }
    w[4] = W(1.0 - s);
  } else {  // ORDER == 5, t in [0, 1), distances 2+t .. 3-t
    double s = 0.0;
    for (int j = 0; j < 5; ++j) {
      const double d = fabs(t + 2.0 - j);
      double v;
      if (d < 1.0) {
        const double d2 = d * d;
        v = d2 * (d2 * (0.25 - d / 12.0) - 0.5) + 0.55;
      } else if (d < 2.0) {
        v = d * (d * (d * (d * (d / 24.0 - 0.375) + 1.25) - 1.75) + 0.625) +
            0.425;
      } else {
        const double e = 3.0 - d;
        v = e * e * e * e * e / 120.0;
      }
      w[j] = W(v);
      s += v;
    }
    w[5] = W(1.0 - s);
  }
}

// Mirror-boundary B-spline prefilter of one line (elements p[i * es],
// n >= 2): gain, then per pole the causal pass with the exact mirror
// initial value (sum over the whole line) and the anti-causal pass
// (Unser/Thevenaz recursive filter), in float64 with values stored as R.
template <typename R>
__device__ __forceinline__ void iir_line(R *p, int n, long long es, int npoles,
                                         double z1, double z2, double gain) {
  for (int i = 0; i < n; ++i) p[i * es] = (R)((double)p[i * es] * gain);
  for (int pi = 0; pi < npoles; ++pi) {
    const double z = pi == 0 ? z1 : z2;
    const double zn1 = pow(z, (double)(n - 1));
    double c0 = (double)p[0] + zn1 * (double)p[(n - 1) * es];
    double zi = z;
    for (int i = 1; i < n - 1; ++i) {
      c0 += zi * ((double)p[i * es] + zn1 * (double)p[(n - 1 - i) * es]);
      zi *= z;
    }
    double prev = c0 / (1.0 - zn1 * zn1);
    p[0] = (R)prev;
    for (int i = 1; i < n; ++i) {
      prev = (double)p[i * es] + z * prev;
      p[i * es] = (R)prev;
    }
    double next = (z * (double)p[(n - 2) * es] + prev) * z / (z * z - 1.0);
    p[(n - 1) * es] = (R)next;
    for (int i = n - 2; i >= 0; --i) {
      next = z * (next - (double)p[i * es]);
      p[i * es] = (R)next;
    }
  }
}

// Centred B-spline of an order (2..5) at distance d: the weight the
// forward interpolation gives the tap at distance d (zero outside
// |d| < (ORDER + 1) / 2).
template <int ORDER>
__device__ __forceinline__ double bspline_val(double d) {
  d = fabs(d);
  if (ORDER == 2) {
    if (d < 0.5) return 0.75 - d * d;
    if (d < 1.5) return 0.5 * (1.5 - d) * (1.5 - d);
    return 0.0;
  } else if (ORDER == 3) {
    if (d < 1.0) return (d * d * (d - 2.0) * 3.0 + 4.0) / 6.0;
    if (d < 2.0) return (2.0 - d) * (2.0 - d) * (2.0 - d) / 6.0;
    return 0.0;
  } else if (ORDER == 4) {
    if (d < 0.5) {
      const double d2 = d * d;
      return d2 * (d2 * 0.25 - 0.625) + 115.0 / 192.0;
    }
    if (d < 1.5)
      return (55.0 + d * (20.0 + d * (-120.0 + d * (80.0 - 16.0 * d)))) /
             96.0;
    if (d < 2.5) {
      const double e = 2.5 - d;
      return e * e * e * e / 24.0;
    }
    return 0.0;
  } else {
    if (d < 1.0) {
      const double d2 = d * d;
      return d2 * (d2 * (0.25 - d / 12.0) - 0.5) + 0.55;
    }
    if (d < 2.0)
      return d * (d * (d * (d * (d / 24.0 - 0.375) + 1.25) - 1.75) + 0.625) +
             0.425;
    if (d < 3.0) {
      const double e = 3.0 - d;
      return e * e * e * e * e / 120.0;
    }
    return 0.0;
  }
}

// Same weights from the fractional offset t = c - floor(c) (odd orders) or
// t = c - floor(c + 0.5) (even orders), evaluated in W: the float32 path
// forms t exactly in float64 and only the polynomials run in float32.
// x / d for the float32 weights as a product with the rounded reciprocal
// (an IEEE division costs about seven VALU instructions, and the gather
// kernels form these weights per voxel and angle); float64 divides.
template <typename W>
__device__ __forceinline__ W div_const(W x, double d) {
  if constexpr (sizeof(W) == 4)
    return x * (W)(1.0 / d);
  else
    return x / (W)d;
}

template <int ORDER, typename W>
__device__ __forceinline__ void spline_weights_t(W t, W *w) {
  if (ORDER == 0) {
    w[0] = W(1);
  } else if (ORDER == 1) {
    w[0] = W(1) - t;
    w[1] = t;
  } else if (ORDER == 2) {
    const W a = W(0.5) - t;
    w[0] = W(0.5) * a * a;
    w[1] = W(0.75) - t * t;
    w[2] = W(1) - w[0] - w[1];
  } else if (ORDER == 3) {
    const W z = W(1) - t;
    w[0] = div_const(z * z * z, 6.0);
    w[1] = div_const(t * t * (t - W(2)) * W(3) + W(4), 6.0);
    w[2] = div_const(z * z * (z - W(2)) * W(3) + W(4), 6.0);
    w[3] = W(1) - w[0] - w[1] - w[2];
  } else if (ORDER == 4) {
    // distances 2+t, 1+t, |t|, 1-t (t in [-0.5, 0.5)); the last weight
    // closes the partition of unity
    const W d0 = W(2) + t, d1 = W(1) + t, d3 = W(1) - t;
    const W ta = t < W(0) ? -t : t;
    const W e0 = W(2.5) - d0;
    w[0] = div_const(e0 * e0 * e0 * e0, 24.0);
    w[1] = div_const(W(55) + d1 * (W(20) + d1 * (W(-120) + d1 * (W(80) - W(16) * d1))), 96.0);
    const W t2 = ta * ta;
    w[2] = t2 * (t2 * W(0.25) - W(0.625)) + W(115) / W(192);
    w[3] = div_const(W(55) + d3 * (W(20) + d3 * (W(-120) + d3 * (W(80) - W(16) * d3))), 96.0);
    w[4] = W(1) - w[0] - w[1] - w[2] - w[3];
  } else {  // ORDER == 5, t in [0, 1): distances 2+t, 1+t, t, 1-t, 2-t
    const W d[5] = {W(2) + t, W(1) + t, t, W(1) - t, W(2) - t};
    W s = W(0);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const W x = d[j];
      W v;
      if (j == 0) {  // [2, 3)
        const W e = W(3) - x;
        v = div_const(e * e * e * e * e, 120.0);
      } else if (j == 2 || j == 3) {  // [0, 1]
        const W x2 = x * x;
        v = x2 * (x2 * (W(0.25) - div_const(x, 12.0)) - W(0.5)) + W(0.55);
      } else {  // [1, 2]
        v = x * (x * (x * (x * (div_const(x, 24.0) - W(0.375)) + W(1.25)) - W(1.75)) +
                 W(0.625)) + W(0.425);
      }
      w[j] = v;
      s += v;
    }
    w[5] = W(1) - s;
  }
}

// First tap of the spline at coordinate c (float64) and its weights in R:
// float32 weights come from the exact float64 fractional offset, float64
// weights from spline_weights.
template <int ORDER, typename R>
__device__ __forceinline__ int spline_taps(double c, R *w) {
  const double f = (ORDER & 1) ? floor(c) : floor(c + 0.5);
  if constexpr (sizeof(R) == 4)
    spline_weights_t<ORDER, R>((R)(c - f), w);
  else
    spline_weights<ORDER, R>(c, w);
  return (int)f - ORDER / 2;
}

// Input coordinate of an affine map with scipy.ndimage's operation order:
// ((0 + o0*m0) + o1*m1 [+ o2*m2]) + off, no fused multiply-add.
__device__ __forceinline__ double affine2(double o0, double o1, double m0,
                                          double m1, double off) {
#pragma clang fp contract(off)
  double t = o0 * m0;
  t = t + o1 * m1;
  return t + off;
}
__device__ __forceinline__ double affine3(double o0, double o1, double o2,
                                          double m0, double m1, double m2,
                                          double off) {
#pragma clang fp contract(off)
  double t = o0 * m0;
  t = t + o1 * m1;
  t = t + o2 * m2;
  return t + off;
}

// ---------------------------------------------------------------------------
// propagator exp(i*phi), phi = km*(M-1)*z computed in float64.
// Float32: the phase is reduced in float64 to quarter turns q and a
// remainder x in [-pi/4, pi/4] (exact enough: |phi| < 2^20 here), then sin x
// and cos x are Taylor polynomials through x^9 and x^8 (truncation below
// 2e-9 on that interval), evaluated in float32 with FMA, and rotated by q
// quarter turns (exact).  Max abs error below 1e-7 (4e-8 for sin, 9e-8
// for cos when evaluated without FMA), against 1.3e-7 for the hardware
// sin/cos; the float32 path's accuracy depends on it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ cpx<float> phasor(double phi, float) {
  const double qt = phi * 0.63661977236758134;  // phi / (pi/2)
  const double qn = rint(qt);
  const float x = (float)((qt - qn) * 1.5707963267948966);
  const float x2 = x * x;
  const float sn =
      x + x * x2 *
              (-1.6666667e-01f +
               x2 * (8.3333333e-03f + x2 * (-1.9841270e-04f + x2 * 2.7557319e-06f)));
  const float cs =
      1.0f + x2 * (-0.5f + x2 * (4.1666668e-02f +
                                 x2 * (-1.3888889e-03f + x2 * 2.4801587e-05f)));
  const int q = (int)((long long)qn & 3);
  return q == 0 ? mk<float>(cs, sn)
                : q == 1 ? mk<float>(-sn, cs)
                         : q == 2 ? mk<float>(-cs, -sn) : mk<float>(sn, -cs);
}
__device__ __forceinline__ cpx<double> phasor(double phi, double) {
  double s, c;
  sincos(phi, &s, &c);
  return mk<double>(c, s);
}


}  // namespace odt

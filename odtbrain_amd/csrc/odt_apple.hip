// Missing-apple-core correction (odtbrain/apple.py:124-262): iterative
// projections between the object domain (constraints) and the Fourier
// domain (measured frequencies outside the apple core restored, envelope
// suppression inside it), computed in complex128 on the volume.
//
//   k_apple_core      the core mask (apple.py:7-49, numpy's operation order)
//   k_apple_init      data = Re f
//   k_transpose3      batched 2D transposes through LDS tiles: with the
//                     batched 1D transforms they make the 3D transform
//   k_env_*           envelope (apple.py:288-353): |F| outside the core,
//                     fftshift into a cube, core filled from the (1,0,2)
//                     transpose, leftovers from the nearest valid sample,
//                     separable Gaussian (truncate 4, zero outside)
//   k_constraint_nn   apple.py:52-71
//   k_sh_*            symmetric-histogram constraint, apple.py:74-121: range
//                     of Re f over the spill-over region (and background
//                     shell), numpy's 100-bin histogram of Re f, the
//                     half-maximum band around its peak, values outside
//                     the band halved
//   k_gauss_weights   scipy.ndimage's normalised Gaussian taps
//   k_apple_fourier   envelope suppression, measured data outside the core,
//                     per-block sums of |F| over the core
//   k_scale           1/N after the inverse transform
#include <hip/hip_runtime.h>

#include "odt_device.h"
#include "odt_launch.h"

namespace odt {

namespace {
constexpr double kPiA = 3.141592653589793;
__device__ __forceinline__ double fftfreq(int i, int n) {
  // numpy.fft.fftfreq: i / n for i < (n + 1) / 2, (i - n) / n after,
  // formed as an integer times 1 / n
  const int j = i < (n - 1) / 2 + 1 ? i : i - n;
  return (double)j * (1.0 / (double)n);
}
}  // namespace

__global__ void __launch_bounds__(256)
    k_apple_core(unsigned char *core, int n0, int n1, int n2, double km) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
  const double fx = fftfreq(a, n0), fy = fftfreq(b, n1), fz = fftfreq(c, n2);
  const double twopi = 2 * kPiA;
  const double ky = twopi * fy;
  const double kxz = twopi * sqrt(fx * fx + fz * fz);
  const double kr = twopi * sqrt(fx * fx + fy * fy + fz * fz);
  double root = 2 * km * kxz - kxz * kxz;
  if (root < 0) root = 0;
  bool in = fabs(ky) > sqrt(root);
  if (kr > sqrt(2.0) * km) in = false;
  core[i] = in;
}

template <typename RI>
__global__ void __launch_bounds__(256)
    k_apple_init(const cpx<RI> *f, cpx<double> *data, long long N) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) data[i] = mk<double>((double)f[i].re, 0.0);
}

// out[b][c][r] = in[b][r][c] for batch b of rows x cols matrices
__global__ void __launch_bounds__(256)
    k_transpose3(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ out,
                 int rows, int cols) {
  __shared__ cpx<double> t[32][33];
  const long long b = blockIdx.z;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const cpx<double> *src = in + b * (long long)rows * cols;
  cpx<double> *dst = out + b * (long long)rows * cols;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) t[k][tx] = src[(long long)r * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) dst[(long long)c * rows + r] = t[tx][k];
  }
}

// numpy fftshift source index along an axis of length n
__device__ __forceinline__ int shift_src(int i, int n) { return (i - n / 2 + n) % n; }
// numpy ifftshift source index
__device__ __forceinline__ int ishift_src(int i, int n) { return (i + n / 2) % n; }

// hull_pad (m^3) = fftshift(|F| with NaN in the core) placed at (s0, s1, s2);
// core_pad likewise; zero / false elsewhere
__global__ void __launch_bounds__(256)
    k_env_pad(const cpx<double> *__restrict__ F, const unsigned char *__restrict__ core,
              int n0, int n1, int n2, int m, int s0, int s1, int s2,
              double *__restrict__ hp, unsigned char *__restrict__ cp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long M = (long long)m * m * m;
  if (i >= M) return;
  const int c = (int)(i % m), b = (int)((i / m) % m), a = (int)(i / ((long long)m * m));
  const int a0 = a - s0, b0 = b - s1, c0 = c - s2;
  if (a0 < 0 || a0 >= n0 || b0 < 0 || b0 >= n1 || c0 < 0 || c0 >= n2) {
    hp[i] = 0.0;
    cp[i] = 0;
    return;
  }
  const long long src = ((long long)shift_src(a0, n0) * n1 + shift_src(b0, n1)) * n2 +
                        shift_src(c0, n2);
  const bool in = core[src] != 0;
  const cpx<double> v = F[src];
  hp[i] = in ? (double)NAN : hypot(v.re, v.im);
  cp[i] = in;
}

// hull_pad[core_pad] = transpose(hull_pad, (1, 0, 2))[core_pad] (read from
// the unfilled array); counts the NaNs left
__global__ void __launch_bounds__(256)
    k_env_transfill(const double *__restrict__ hp, const unsigned char *__restrict__ cp,
                    int m, double *__restrict__ out, unsigned *__restrict__ nnan) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long M = (long long)m * m * m;
  if (i >= M) return;
  double v = hp[i];
  if (cp[i]) {
    const int c = (int)(i % m), b = (int)((i / m) % m), a = (int)(i / ((long long)m * m));
    v = hp[((long long)b * m + a) * m + c];
  }
  out[i] = v;
  if (isnan(v)) atomicAdd(nnan, 1u);
}

// NaN samples take the value of the nearest non-NaN sample (Euclidean; on
// ties the smallest linear index), by exhaustive search: only the few
// samples of non-cubic volumes where the core meets its transpose
__global__ void __launch_bounds__(256)
    k_env_nearest(const double *__restrict__ in, int m, double *__restrict__ out) {
  const long long i = blockIdx.x;  // one block per sample
  const long long M = (long long)m * m * m;
  if (i >= M) return;
  const double v = in[i];
  if (!isnan(v)) {
    if (threadIdx.x == 0) out[i] = v;
    return;
  }
  const int c = (int)(i % m), b = (int)((i / m) % m), a = (int)(i / ((long long)m * m));
  long long best = -1, bd = 0x7fffffffffffffffLL;
  for (long long j = threadIdx.x; j < M; j += blockDim.x) {
    if (isnan(in[j])) continue;
    const long long jc = j % m, jb = (j / m) % m, ja = j / ((long long)m * m);
    const long long d = (ja - a) * (ja - a) + (jb - b) * (jb - b) + (jc - c) * (jc - c);
    if (d < bd || (d == bd && j < best)) {
      bd = d;
      best = j;
    }
  }
  __shared__ long long sd[256], sj[256];
  sd[threadIdx.x] = bd;
  sj[threadIdx.x] = best;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const long long od = sd[threadIdx.x + s], oj = sj[threadIdx.x + s];
      if (oj >= 0 && (od < sd[threadIdx.x] || (od == sd[threadIdx.x] && oj < sj[threadIdx.x]))) {
        sd[threadIdx.x] = od;
        sj[threadIdx.x] = oj;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sj[0] >= 0 ? in[sj[0]] : 0.0;
}

// g = fftshift(hull) with hull = ifftshift(hull_pad[slice]): the slice of
// the padded cube itself (shifts cancel), laid out as n0 x n1 x n2
__global__ void __launch_bounds__(256)
    k_env_unpad(const double *__restrict__ hp, int n0, int n1, int n2, int m, int s0,
                int s1, int s2, double *__restrict__ g) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
  g[i] = hp[((long long)(a + s0) * m + (b + s1)) * m + (c + s2)];
}

// one axis of scipy.ndimage.gaussian_filter (correlate1d, mode constant 0)
__global__ void __launch_bounds__(256)
    k_env_gauss(const double *__restrict__ in, double *__restrict__ out, int n0, int n1,
                int n2, int axis, const double *__restrict__ w, int r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
  const int pos = axis == 0 ? a : axis == 1 ? b : c;
  const int n = axis == 0 ? n0 : axis == 1 ? n1 : n2;
  const long long st = axis == 0 ? (long long)n1 * n2 : axis == 1 ? n2 : 1;
  double s = 0.0;
  for (int k = -r; k <= r; ++k) {
    const int q = pos + k;
    if (q >= 0 && q < n) s += w[k + r] * in[i + (long long)k * st];
  }
  out[i] = s;
}

// envelope = ifftshift(g)
__global__ void __launch_bounds__(256)
    k_env_ishift(const double *__restrict__ g, int n0, int n1, int n2,
                 double *__restrict__ env) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
  env[i] = g[((long long)ishift_src(a, n0) * n1 + ishift_src(b, n1)) * n2 + ishift_src(c, n2)];
}

__device__ __forceinline__ double spill_q(int a, int b, int c, double cx, double cy,
                                          double cz, double shell) {
#pragma clang fp contract(off)
  const double u = ((double)a - cx) / (cx - shell), v = ((double)b - cy) / (cy - shell),
               t = ((double)c - cz) / (cz - shell);
  return (u * u + v * v) + t * t;
}

// non-negativity (apple.py:52-71): imaginary part 0, background shell and
// the region outside the inscribed ellipsoid 0, negative real parts 0
// (where the optional mask allows)
__global__ void __launch_bounds__(256)
    k_constraint_nn(cpx<double> *data, int n0, int n1, int n2,
                    const unsigned char *__restrict__ mask, double bg_width) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
  const double cx = n0 / 2.0, cy = n1 / 2.0, cz = n2 / 2.0;
  double re = data[i].re;
  const bool outer = spill_q(a, b, c, cx, cy, cz, 0.0) > 1;
  if (bg_width >= 0) {
    const bool inner = spill_q(a, b, c, cx, cy, cz, bg_width) > 1;
    if (outer != inner) re = 0.0;
  }
  if (outer) re = 0.0;
  if (re < 0 && (!mask || mask[i])) re = 0.0;
  data[i] = mk<double>(re, 0.0);
}

// Symmetric-histogram constraint (apple.py:74-121), in four steps.
// Step 1: imaginary parts 0; per block the minimum and maximum of Re f over
// the spill-over region united with the background shell (bg_width >= 0).
constexpr int kShBlocks = 1024;
constexpr int kShBins = 100;

__global__ void __launch_bounds__(256)
    k_sh_minmax(cpx<double> *data, int n0, int n1, int n2, double bg_width,
                double *__restrict__ part) {
  const long long N = (long long)n0 * n1 * n2;
  const double cx = n0 / 2.0, cy = n1 / 2.0, cz = n2 / 2.0;
  double lo = INFINITY, hi = -INFINITY;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
    const double re = data[i].re;
    data[i] = mk<double>(re, 0.0);
    bool in = spill_q(a, b, c, cx, cy, cz, 0.0) > 1;
    if (bg_width >= 0) in = in || spill_q(a, b, c, cx, cy, cz, bg_width) > 1;
    if (in) {
      lo = fmin(lo, re);
      hi = fmax(hi, re);
    }
  }
  __shared__ double sl[256], sh[256];
  sl[threadIdx.x] = lo;
  sh[threadIdx.x] = hi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + k]);
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sl[0];
    part[2 * blockIdx.x + 1] = sh[0];
  }
}

// Step 2 (one block): fmin, fmax into mm[0..1]; the histogram's outer edges
// into mm[2..3] (numpy widens an empty range by 0.5 on either side); the
// bin counts cleared.
__global__ void __launch_bounds__(256)
    k_sh_range(const double *__restrict__ part, int nparts, double *__restrict__ mm,
               unsigned *__restrict__ hist) {
  __shared__ double sl[256], sh[256];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    lo = fmin(lo, part[2 * i]);
    hi = fmax(hi, part[2 * i + 1]);
  }
  sl[threadIdx.x] = lo;
  sh[threadIdx.x] = hi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + k]);
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x < kShBins) hist[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    mm[0] = sl[0];
    mm[1] = sh[0];
    const bool same = sl[0] == sh[0];
    mm[2] = same ? sl[0] - 0.5 : sl[0];
    mm[3] = same ? sh[0] + 0.5 : sh[0];
  }
}

// Bin edge i of numpy.linspace(first, last, 101): i * step + first, the last
// edge exactly `last`.
__device__ __forceinline__ double sh_edge(int i, double first, double last) {
#pragma clang fp contract(off)
  if (i >= kShBins) return last;
  const double step = (last - first) / (double)kShBins;
  return (double)i * step + first;
}

// Step 3: numpy.histogram(Re f, bins=100, range=(first, last)) with numpy's
// uniform-bin index rule (index from the scaled offset, then corrected
// against the edges; the last bin includes the right edge).
__global__ void __launch_bounds__(256)
    k_sh_hist(const cpx<double> *__restrict__ data, long long N,
              const double *__restrict__ mm, unsigned *__restrict__ hist) {
#pragma clang fp contract(off)
  __shared__ unsigned lh[kShBins];
  for (int i = threadIdx.x; i < kShBins; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const double first = mm[2], last = mm[3];
  const double denom = last - first;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = data[i].re;
    if (!(v >= first && v <= last)) continue;
    long long idx = (long long)(((v - first) / denom) * (double)kShBins);
    if (idx == kShBins) idx -= 1;
    if (v < sh_edge((int)idx, first, last)) idx -= 1;
    if (idx != kShBins - 1 && v >= sh_edge((int)idx + 1, first, last)) idx += 1;
    atomicAdd(&lh[idx], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kShBins; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// Step 4a (one thread): the histogram's peak (first maximum), the band of
// bins down to half of it on either side as apple.py:97-99 and count_to_half
// (apple.py:265-277) define it -- the lower side measured against the bin
// below the peak -- and the bin centres at its ends into mm[4..5].
__global__ void k_sh_band(const unsigned *__restrict__ hist, double *__restrict__ mm) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int c = 0;
  for (int i = 1; i < kShBins; ++i)
    if (hist[i] > hist[c]) c = i;
  int nlo = 0;
  if (c > 0) {
    const double half = (double)hist[c - 1] / 2.0;
    for (int k = c - 2; k >= 0; --k) {
      if ((double)hist[k] < half) break;
      ++nlo;
    }
  }
  int nhi = 0;
  {
    const double half = (double)hist[c] / 2.0;
    for (int k = c + 1; k < kShBins; ++k) {
      if ((double)hist[k] < half) break;
      ++nhi;
    }
  }
  const double first = mm[2], last = mm[3];
  const double de = sh_edge(1, first, last) - sh_edge(0, first, last);
  mm[4] = sh_edge(c - nlo + 1, first, last) - de / 2;
  mm[5] = sh_edge(c + nhi + 1, first, last) - de / 2;
}

// Step 4b: values strictly between fmin and the band's lower centre or
// between its upper centre and fmax are halved (where the mask allows),
// then the background shell is halved.
__global__ void __launch_bounds__(256)
    k_sh_apply(cpx<double> *data, int n0, int n1, int n2,
               const unsigned char *__restrict__ mask, double bg_width,
               const double *__restrict__ mm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)n0 * n1 * n2;
  if (i >= N) return;
  const double fmin_ = mm[0], fmax_ = mm[1], lo = mm[4], hi = mm[5];
  double re = data[i].re;
  bool out = (re > fmin_ && re < lo) || (re > hi && re < fmax_);
  if (mask && !mask[i]) out = false;
  if (out) re *= 0.5;
  if (bg_width >= 0) {
    const int c = (int)(i % n2), b = (int)((i / n2) % n1), a = (int)(i / ((long long)n1 * n2));
    const double cx = n0 / 2.0, cy = n1 / 2.0, cz = n2 / 2.0;
    const bool outer = spill_q(a, b, c, cx, cy, cz, 0.0) > 1;
    const bool inner = spill_q(a, b, c, cx, cy, cz, bg_width) > 1;
    if (outer != inner) re *= 0.5;
  }
  data[i] = mk<double>(re, 0.0);
}

// scipy.ndimage's Gaussian taps (_gaussian_kernel1d: exp(-x^2 / (2 sigma^2))
// for x = -r..r, divided by their sum), one thread, in order.
__global__ void k_gauss_weights(double *w, int r, double sigma) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0;
  for (int k = -r; k <= r; ++k) {
    const double v = exp(-0.5 / (sigma * sigma) * (double)k * (double)k);
    w[k + r] = v;
    s += v;
  }
  for (int k = 0; k <= 2 * r; ++k) w[k] /= s;
}

// Fourier step: suppression above the envelope, the measured data outside
// the core, per-block sums of |F| over the core.  sh: the zero frequency of
// the measured data becomes the mean of the two first.
__global__ void __launch_bounds__(256)
    k_apple_fourier(cpx<double> *F, cpx<double> *F0, const double *__restrict__ env,
                    const unsigned char *__restrict__ core, long long N, double factor,
                    int sh, double *__restrict__ part) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < N) {
    cpx<double> v = F[i];
    if (env && hypot(v.re, v.im) > env[i]) v = mk<double>(v.re * factor, v.im * factor);
    if (sh && i == 0) F0[0] = mk<double>((F0[0].re + v.re) / 2, (F0[0].im + v.im) / 2);
    if (!core[i]) v = F0[i];
    else s = hypot(v.re, v.im);
    F[i] = v;
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(256)
    k_sum_parts(const double *__restrict__ part, long long n, double *__restrict__ out) {
  __shared__ double sm[256];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

__global__ void __launch_bounds__(256)
    k_scale_out(const cpx<double> *__restrict__ in, long long N, double s, int out64,
                void *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const cpx<double> v = in[i];
  if (out64)
    reinterpret_cast<cpx<double> *>(out)[i] = mk<double>(v.re * s, v.im * s);
  else
    reinterpret_cast<cpx<float> *>(out)[i] = mk<float>((float)(v.re * s), (float)(v.im * s));
}

namespace {
unsigned nb256(long long n) { return (unsigned)((n + 255) / 256); }
}  // namespace

hipError_t launch_apple_core(unsigned char *core, int n0, int n1, int n2, double km,
                             hipStream_t st) {
  k_apple_core<<<nb256((long long)n0 * n1 * n2), 256, 0, st>>>(core, n0, n1, n2, km);
  return hipGetLastError();
}

hipError_t launch_apple_init(const void *f, int f64, cpx<double> *data, long long N,
                             hipStream_t st) {
  if (f64)
    k_apple_init<double><<<nb256(N), 256, 0, st>>>((const cpx<double> *)f, data, N);
  else
    k_apple_init<float><<<nb256(N), 256, 0, st>>>((const cpx<float> *)f, data, N);
  return hipGetLastError();
}

hipError_t launch_transpose3(const cpx<double> *in, cpx<double> *out, long long batch,
                             int rows, int cols, hipStream_t st) {
  if (batch > 65535) return hipErrorInvalidValue;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, (unsigned)batch);
  k_transpose3<<<grid, 256, 0, st>>>(in, out, rows, cols);
  return hipGetLastError();
}

hipError_t launch_envelope(const cpx<double> *F, const unsigned char *core, int n0, int n1,
                           int n2, double *hp, double *hp2, unsigned char *cp,
                           unsigned *nnan_dev, const double *w, int r, double *g, double *g2,
                           double *env, hipStream_t st) {
  const int m = n0 > n1 ? n0 : n1;
  const int dsh = (n1 > n0 ? n1 - n0 : n0 - n1) / 2;
  int s0, s1, s2;
  if (n0 > n1) {
    s0 = 0;
    s1 = dsh;
    s2 = 0;
  } else {
    s0 = dsh;
    s1 = 0;
    s2 = dsh;
  }
  const long long M = (long long)m * m * m, N = (long long)n0 * n1 * n2;
  k_env_pad<<<nb256(M), 256, 0, st>>>(F, core, n0, n1, n2, m, s0, s1, s2, hp, cp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_env_transfill<<<nb256(M), 256, 0, st>>>(hp, cp, m, hp2, nnan_dev);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  unsigned nnan = 0;
  if ((e = hipMemcpyAsync(&nnan, nnan_dev, sizeof(unsigned), hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  const double *filled = hp2;
  if (nnan > 0) {
    if (M > 0x7fffffffLL) return hipErrorInvalidValue;
    k_env_nearest<<<(unsigned)M, 256, 0, st>>>(hp2, m, hp);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    filled = hp;
  }
  k_env_unpad<<<nb256(N), 256, 0, st>>>(filled, n0, n1, n2, m, s0, s1, s2, g);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_env_gauss<<<nb256(N), 256, 0, st>>>(g, g2, n0, n1, n2, 0, w, r);
  k_env_gauss<<<nb256(N), 256, 0, st>>>(g2, g, n0, n1, n2, 1, w, r);
  k_env_gauss<<<nb256(N), 256, 0, st>>>(g, g2, n0, n1, n2, 2, w, r);
  k_env_ishift<<<nb256(N), 256, 0, st>>>(g2, n0, n1, n2, env);
  return hipGetLastError();
}

hipError_t launch_constraint_nn(cpx<double> *data, int n0, int n1, int n2,
                                const unsigned char *mask, double bg_width, hipStream_t st) {
  k_constraint_nn<<<nb256((long long)n0 * n1 * n2), 256, 0, st>>>(data, n0, n1, n2, mask,
                                                                   bg_width);
  return hipGetLastError();
}

hipError_t launch_constraint_sh(cpx<double> *data, int n0, int n1, int n2,
                                const unsigned char *mask, double bg_width, double *part,
                                double *mm, unsigned *hist, hipStream_t st) {
  const long long N = (long long)n0 * n1 * n2;
  k_sh_minmax<<<kShBlocks, 256, 0, st>>>(data, n0, n1, n2, bg_width, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_sh_range<<<1, 256, 0, st>>>(part, kShBlocks, mm, hist);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_sh_hist<<<kShBlocks, 256, 0, st>>>(data, N, mm, hist);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_sh_band<<<1, 64, 0, st>>>(hist, mm);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_sh_apply<<<nb256(N), 256, 0, st>>>(data, n0, n1, n2, mask, bg_width, mm);
  return hipGetLastError();
}

size_t sh_part_bytes() { return (size_t)2 * kShBlocks * sizeof(double); }
size_t sh_hist_bytes() { return (size_t)kShBins * sizeof(unsigned); }

hipError_t launch_gauss_weights(double *w, int r, double sigma, hipStream_t st) {
  k_gauss_weights<<<1, 64, 0, st>>>(w, r, sigma);
  return hipGetLastError();
}

hipError_t launch_apple_fourier(cpx<double> *F, cpx<double> *F0, const double *env,
                                const unsigned char *core, long long N, double factor,
                                int sh, double *part, double *sum, hipStream_t st) {
  const unsigned nb = nb256(N);
  k_apple_fourier<<<nb, 256, 0, st>>>(F, F0, env, core, N, factor, sh, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_sum_parts<<<1, 256, 0, st>>>(part, nb, sum);
  return hipGetLastError();
}

hipError_t launch_scale_out(const cpx<double> *in, long long N, double s, int out64,
                            void *out, hipStream_t st) {
  k_scale_out<<<nb256(N), 256, 0, st>>>(in, N, s, out64, out);
  return hipGetLastError();
}

}  // namespace odt

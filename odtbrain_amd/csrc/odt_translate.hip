// Elementwise translation of a reconstructed object function to refractive
// index (odtbrain/_translate_ri.py): odt_to_ri (:5-57) and opt_to_ri
// (:60-90).  HBM-bound streams, 16 bytes per lane and iteration for complex
// float32 (two elements), grid-stride over the array.
#include <hip/hip_runtime.h>

#include "odt_device.h"
#include "odt_launch.h"

namespace odt {

// Principal square root of x + iy (non-negative real part, the sign of the
// imaginary part follows y, -0 included), as C99 csqrt / numpy compute it:
// the larger component from sqrt((|x| + |z|) / 2), the other by division.
template <typename R>
__device__ __forceinline__ cpx<R> csqrt_principal(R x, R y) {
  const R h = hypot(x, y);
  if (h == R(0)) return mk<R>(R(0), y);
  if (!(h < INFINITY)) {  // infinite or NaN components
    if (isinf(y)) return mk<R>(R(INFINITY), y);
    if (isnan(x) || isnan(y)) return mk<R>(R(NAN), R(NAN));
    return x > 0 ? mk<R>(x, copysign(R(0), y)) : mk<R>(R(0), copysign(R(INFINITY), y));
  }
  if (x >= R(0)) {
    const R t = sqrt(R(0.5) * (h + x));
    return mk<R>(t, R(0.5) * (y / t));
  }
  const R t = sqrt(R(0.5) * (h - x));
  return mk<R>(fabs(R(0.5) * (y / t)), copysign(t, y));
}

// ri = nm * sqrt(f / km2 + 1); a root with negative real part is negated
// (never the case for the principal root; NaN passes through).
template <typename R>
__global__ void __launch_bounds__(256)
    k_odt_to_ri_c(const cpx<R> *__restrict__ f, cpx<R> *__restrict__ ri,
                  long long n, R km2, R nm) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const cpx<R> v = f[i];
    // the radicand is f / km2 + (1 + 0i): the imaginary part gains +0, so
    // a -0 imaginary part becomes +0 before the root, as in numpy
    cpx<R> r = csqrt_principal<R>(v.re / km2 + R(1), v.im / km2 + R(0));
    r = mk<R>(nm * r.re, nm * r.im);
    if (r.re < R(0)) r = mk<R>(-r.re, -r.im);
    ri[i] = r;
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_odt_to_ri_r(const R *__restrict__ f, R *__restrict__ ri, long long n,
                  R km2, R nm) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    R r = nm * sqrt(f[i] / km2 + R(1));
    if (r < R(0)) r = -r;
    ri[i] = r;
  }
}

// ri = nm + f / (2 pi) * res (complex f: the real part is shifted by nm)
template <typename R>
__global__ void __launch_bounds__(256)
    k_opt_to_ri(const R *__restrict__ f, R *__restrict__ ri, long long n,
                int is_complex, R twopi, R res, R nm) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const R v = f[i] / twopi * res;
    ri[i] = (!is_complex || (i & 1) == 0) ? nm + v : v;
  }
}

static unsigned ew_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

template <typename R>
hipError_t launch_odt_to_ri(int is_complex, long long n, double km2, double nm,
                            const void *f, void *ri, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (is_complex)
    k_odt_to_ri_c<R><<<ew_blocks(n), 256, 0, st>>>(
        (const cpx<R> *)f, (cpx<R> *)ri, n, (R)km2, (R)nm);
  else
    k_odt_to_ri_r<R><<<ew_blocks(n), 256, 0, st>>>((const R *)f, (R *)ri, n,
                                                    (R)km2, (R)nm);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_opt_to_ri(int is_complex, long long n, double res, double nm,
                            const void *f, void *ri, hipStream_t st) {
  const long long nr = is_complex ? 2 * n : n;
  if (nr <= 0) return hipSuccess;
  k_opt_to_ri<R><<<ew_blocks(nr), 256, 0, st>>>(
      (const R *)f, (R *)ri, nr, is_complex, (R)(2.0 * M_PI), (R)res, (R)nm);
  return hipGetLastError();
}

template hipError_t launch_odt_to_ri<float>(int, long long, double, double,
                                            const void *, void *, hipStream_t);
template hipError_t launch_odt_to_ri<double>(int, long long, double, double,
                                             const void *, void *, hipStream_t);
template hipError_t launch_opt_to_ri<float>(int, long long, double, double,
                                            const void *, void *, hipStream_t);
template hipError_t launch_opt_to_ri<double>(int, long long, double, double,
                                             const void *, void *, hipStream_t);

}  // namespace odt

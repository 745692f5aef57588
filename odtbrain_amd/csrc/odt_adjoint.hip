// Adjoint of the filtered backpropagation about the y axis.
//
// The reconstruction of backpropagate_3d (odtbrain/_alg3d_bpp.py:493-546) is
// linear in the sinogram for edge or no padding.  With the explicit
// prefilter it reads, per angle a,
//   f += R_a Q_z Q_x Crop IDFT_x IDFT_y (e^{i k_z z} P) DFT_y DFT_x Pad (w_a u_a)
// (DFT unnormalised e^{-i}, IDFT unnormalised e^{+i}, P the prefactor table
// restricted to the transform support, Q the mirror B-spline prefilter, R_a
// the rotate-and-sum of scipy.ndimage.rotate).  Its adjoint, applied to an
// output gradient g, is
//   u_a = w_a Pad^T DFT_x^H DFT_y^H conj(P) sum_z e^{-i k_z z}
//         DFT_y IDFT_x^H ... = see the kernels below, stage by stage:
//   k_adj_rotate     R_a^T g: the rotate gather transposed into a scatter
//                    (same coordinates, weights and mirrored taps; atomics)
//   k_scale_ends +   Q^T = D Q D^{-1} per axis, D = diag(1/2, 1, ..., 1, 1/2)
//   k_iir            (whole-sample mirror extension; checked numerically)
//   k_adj_row        forward transform along x of each zero-padded slab row,
//                    support columns kept (IDFT_x^H = DFT_x)
//   k_adj_col        forward transform along y of each support column and
//                    each z, times e^{-i k_z z}, summed over z in registers
//   k_adj_inv_col    times conj(P), inverse transform along y, Pad^T in y
//   k_adj_inv_row    inverse transform along x, Pad^T in x, times w_a
// Transforms use the register radix-8 engine for lengths 64..512 (one
// transform per wavefront or less).
#include <hip/hip_runtime.h>

#include "odt_device.h"
#include "odt_fft8.h"
#include "odt_launch.h"

namespace odt {

extern __shared__ __align__(16) unsigned char odt_smem[];

template <typename R>
__global__ void k_fill_zero(R *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = R(0);
}

// R_a^T: every output voxel (z, y, x) of every angle adds g * w to the taps
// it reads in the forward gather (k_rotate_lds / scipy's rotate).
template <typename R, int ORDER>
__global__ void __launch_bounds__(256)
    k_adj_rotate(const cpx<R> *__restrict__ g, const double *__restrict__ params,
                 cpx<R> *__restrict__ slab, int ny, int nz, int nx) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= nz * nx) return;
  const int z = ix / nx, x = ix - z * nx;
  const int y = blockIdx.y, a = blockIdx.z;
  const double *p = params + 6 * (size_t)a;
  const double cz = affine2((double)z, (double)x, p[0], p[1], p[4]);
  const double cx = affine2((double)z, (double)x, p[2], p[3], p[5]);
  if (cz < 0.0 || cz > (double)(nz - 1) || cx < 0.0 || cx > (double)(nx - 1))
    return;
  double wz[ORDER + 1], wx[ORDER + 1];
  spline_weights<ORDER, double>(cz, wz);
  spline_weights<ORDER, double>(cx, wx);
  const int sz = spline_start<ORDER>(cz), sx = spline_start<ORDER>(cx);
  const cpx<R> gv = g[((size_t)z * ny + y) * nx + x];
  R *pl = reinterpret_cast<R *>(slab + ((size_t)a * ny + y) * nz * nx);
#pragma unroll
  for (int i = 0; i <= ORDER; ++i) {
    const int zr = mirror_idx(sz + i, nz);
#pragma unroll
    for (int j = 0; j <= ORDER; ++j) {
      const int xc = mirror_idx(sx + j, nx);
      const double w = wz[i] * wx[j];
      R *dst = pl + 2 * ((size_t)zr * nx + xc);
      unsafeAtomicAdd(dst, (R)(w * (double)gv.re));
      unsafeAtomicAdd(dst + 1, (R)(w * (double)gv.im));
    }
  }
}

// Multiply both end samples of each line by s (line addressing of k_iir).
template <typename R>
__global__ void k_scale_ends(R *data, long long nlines, int n, long long n0,
                             long long s0, long long s1, long long es, R s) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines || n < 2) return;
  R *p = data + (l / n0) * s1 + (l % n0) * s0;
  p[0] *= s;
  p[(n - 1) * es] *= s;
}

template <int N>
__device__ __forceinline__ int sfreq(int bin) {
  return bin < (N + 1) / 2 ? bin : bin - N;
}

// Forward transform along x of slab rows (a, y, z) placed at padxl in a
// zero row of lNx = N; support columns -> Tp[a][z][y][cx] (stride ldt).
// Block: rows y0..y0+rb-1 of slice k of angle a, tpb rows at a time.
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_row(const cpx<R> *__restrict__ slab, cpx<R> *__restrict__ Tp,
              const cpx<R> *__restrict__ twx, Geom g, int rb, int ldt) {
  using FFT = Fft8<R, N, -1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 64, "one transform per wavefront at most");
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *b0 = reinterpret_cast<cpx<R> *>(odt_smem) + t * FFT::PADN;
  const int y0 = blockIdx.x * rb, k = blockIdx.y, ai = blockIdx.z;
  const int ln = g.ln, ny = g.Ny, nx = g.Nx;
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twx, j);
  for (int r = t; r < rb; r += tpb) {
    const int y = y0 + r;
    if (y >= ny) break;
    const cpx<R> *src = slab + (((size_t)ai * ny + y) * ln + k) * nx;
    cpx<R> v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int x = j + m * TT - g.padxl;
      v[m] = (x >= 0 && x < nx) ? src[x] : mk<R>(0, 0);
    }
    FFT::run(v, b0, nullptr, tw, j);
    cpx<R> *dst = Tp + (((size_t)ai * ln + k) * ny + y) * ldt;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int cx = sfreq<N>(j + m * TT) - g.jx0;
      if (cx >= 0 && cx < g.Lx) dst[cx] = v[m];
    }
  }
}

// Per support column cx of angle a: for every slice z, forward transform
// along y of Tp[a][z][0..ny)[cx] placed at padyl, times e^{-i k_z(c,cy) z},
// summed over z; support rows -> Sp[a][cx][cy].  One column per transform
// slot.
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_col(const cpx<R> *__restrict__ Tp, cpx<R> *__restrict__ Sp,
              const cpx<R> *__restrict__ twy, const double *__restrict__ prop_a,
              Geom g, int ldt) {
  using FFT = Fft8<R, N, -1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 64, "one transform per wavefront at most");
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *b0 = reinterpret_cast<cpx<R> *>(odt_smem) + t * FFT::PADN;
  const int c = blockIdx.x * tpb + t, ai = blockIdx.y;
  const bool live = c < g.Lx;
  const int cs = live ? c : 0;
  const int ln = g.ln, ny = g.Ny;
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twy, j);
  int cyc[8];
  bool cyv[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cy = sfreq<N>(j + m * TT) - g.jy0;
    cyv[m] = cy >= 0 && cy < g.Ly;
    cyc[m] = cyv[m] ? cy : 0;
  }
  const double *pa = prop_a + (size_t)cs * g.Ly;
  cpx<R> acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = mk<R>(0, 0);
  for (int k = 0; k < ln; ++k) {
    const cpx<R> *src = Tp + (((size_t)ai * ln + k) * ny) * ldt + cs;
    cpx<R> v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int y = j + m * TT - g.padyl;
      v[m] = (y >= 0 && y < ny) ? src[(size_t)y * ldt] : mk<R>(0, 0);
    }
    FFT::run(v, b0, nullptr, tw, j);
    const double z = g.z0 + (double)k;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if (!cyv[m]) continue;
      const cpx<R> e = phasor(-pa[cyc[m]] * z, R(0));
      acc[m] = acc[m] + v[m] * e;
    }
  }
  if (!live) return;
  cpx<R> *dst = Sp + ((size_t)ai * g.Lx + c) * g.Ly;
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if (cyv[m]) dst[cyc[m]] = acc[m];
}

// Sum of the padded samples that np.pad copies to detector sample r
// (edge mode: the border samples collect the padding; none: identity).
// buf holds the lN transform outputs of one line in LDS (padded layout).
template <typename R>
__device__ __forceinline__ cpx<R> fold_pad(const cpx<R> *buf, int r, int n,
                                           int padl, int lN) {
  cpx<R> s = buf[lds_pad(padl + r)];
  if (r == 0)
    for (int q = 0; q < padl; ++q) s = s + buf[lds_pad(q)];
  if (r == n - 1)
    for (int q = padl + n; q < lN; ++q) s = s + buf[lds_pad(q)];
  return s;
}

// Per support column cx of angle a: conj(P) * Sp, inverse transform along
// y, padding folded -> U1[a][r][cx] (stride ldt), r in [0, ny).
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_inv_col(const cpx<R> *__restrict__ Sp, const cpx<R> *__restrict__ P,
                  cpx<R> *__restrict__ U1, const cpx<R> *__restrict__ twy,
                  Geom g, int ldt) {
  using FFT = Fft8<R, N, +1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 64, "one transform per wavefront at most");
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *b0 = reinterpret_cast<cpx<R> *>(odt_smem) + t * FFT::PADN;
  const int c = blockIdx.x * tpb + t, ai = blockIdx.y;
  if (c >= g.Lx) return;  // whole slot; transforms are wave-local
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twy, j);
  const cpx<R> *src = Sp + ((size_t)ai * g.Lx + c) * g.Ly;
  const cpx<R> *pc = P + (size_t)c * g.Ly;
  cpx<R> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cy = sfreq<N>(j + m * TT) - g.jy0;
    v[m] = mk<R>(0, 0);
    if (cy >= 0 && cy < g.Ly) {
      const cpx<R> pp = pc[cy];
      v[m] = src[cy] * mk<R>(pp.re, -pp.im);
    }
  }
  FFT::run(v, b0, nullptr, tw, j);
  // natural order into this slot's LDS line, then fold the padding
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
  for (int m = 0; m < 8; ++m) b0[lds_pad(j + m * TT)] = v[m];
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  const int ny = g.Ny;
  for (int r = j; r < ny; r += TT)
    U1[((size_t)ai * ny + r) * ldt + c] = fold_pad(b0, r, ny, g.padyl, N);
}

// Per detector row r of angle a: inverse transform along x of the support
// columns U1[a][r][0..Lx), padding folded, times the angle weight ->
// the sinogram gradient gu[a][r][0..nx).
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_inv_row(const cpx<R> *__restrict__ U1, cpx<R> *__restrict__ gu,
                  const cpx<R> *__restrict__ twx,
                  const double *__restrict__ weights, int a0, Geom g,
                  int ldt) {
  using FFT = Fft8<R, N, +1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 64, "one transform per wavefront at most");
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *b0 = reinterpret_cast<cpx<R> *>(odt_smem) + t * FFT::PADN;
  const int r = blockIdx.x * tpb + t, ai = blockIdx.y;
  const int ny = g.Ny, nx = g.Nx;
  if (r >= ny) return;
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twx, j);
  const cpx<R> *src = U1 + ((size_t)ai * ny + r) * ldt;
  cpx<R> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cx = sfreq<N>(j + m * TT) - g.jx0;
    v[m] = (cx >= 0 && cx < g.Lx) ? src[cx] : mk<R>(0, 0);
  }
  FFT::run(v, b0, nullptr, tw, j);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
  for (int m = 0; m < 8; ++m) b0[lds_pad(j + m * TT)] = v[m];
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  const R w = (R)weights[a0 + ai];
  cpx<R> *dst = gu + ((size_t)(a0 + ai) * ny + r) * nx;
  for (int x = j; x < nx; x += TT) dst[x] = w * fold_pad(b0, x, nx, g.padxl, N);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, int ORDER>
static hipError_t adj_rotate_launch(const cpx<R> *g, int nb, const double *params,
                                    cpx<R> *slab, int ny, int nz, int nx,
                                    hipStream_t st) {
  dim3 grid((nz * nx + 255) / 256, ny, nb);
  k_adj_rotate<R, ORDER><<<grid, 256, 0, st>>>(g, params, slab, ny, nz, nx);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_adj_rotate(int order, const cpx<R> *g, int nb,
                             const double *params, cpx<R> *slab, int ny, int nz,
                             int nx, hipStream_t st) {
  const size_t n = (size_t)nb * ny * nz * nx * 2;
  k_fill_zero<R><<<(unsigned)((n + 255) / 256), 256, 0, st>>>((R *)slab, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (order) {
    case 2: return adj_rotate_launch<R, 2>(g, nb, params, slab, ny, nz, nx, st);
    case 3: return adj_rotate_launch<R, 3>(g, nb, params, slab, ny, nz, nx, st);
    case 4: return adj_rotate_launch<R, 4>(g, nb, params, slab, ny, nz, nx, st);
    case 5: return adj_rotate_launch<R, 5>(g, nb, params, slab, ny, nz, nx, st);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_scale_ends(R *data, long long nlines, int n, long long n0,
                             long long s0, long long s1, long long es, double s,
                             hipStream_t st) {
  if (nlines <= 0) return hipSuccess;
  k_scale_ends<R><<<(unsigned)((nlines + 255) / 256), 256, 0, st>>>(
      data, nlines, n, n0, s0, s1, es, (R)s);
  return hipGetLastError();
}

template <typename R, int N>
static hipError_t adj_fft_launch(int stage, const void *in, void *out,
                                 const void *tw, const void *aux,
                                 const double *prop_a, const double *weights,
                                 int a0, const Geom &g, int na, int ldt,
                                 hipStream_t st) {
  constexpr int TT = N / 8;
  const int tpb = 256 / TT;
  const size_t sh = (size_t)tpb * lds_padded(N) * sizeof(cpx<R>);
  if (stage == 0) {
    const int rb = 16;
    dim3 grid((g.Ny + rb - 1) / rb, g.ln, na);
    k_adj_row<R, N><<<grid, 256, sh, st>>>((const cpx<R> *)in, (cpx<R> *)out,
                                           (const cpx<R> *)tw, g, rb, ldt);
  } else if (stage == 1) {
    dim3 grid((g.Lx + tpb - 1) / tpb, na);
    k_adj_col<R, N><<<grid, 256, sh, st>>>((const cpx<R> *)in, (cpx<R> *)out,
                                           (const cpx<R> *)tw, prop_a, g, ldt);
  } else if (stage == 2) {
    dim3 grid((g.Lx + tpb - 1) / tpb, na);
    k_adj_inv_col<R, N><<<grid, 256, sh, st>>>(
        (const cpx<R> *)in, (const cpx<R> *)aux, (cpx<R> *)out,
        (const cpx<R> *)tw, g, ldt);
  } else {
    dim3 grid((g.Ny + tpb - 1) / tpb, na);
    k_adj_inv_row<R, N><<<grid, 256, sh, st>>>((const cpx<R> *)in, (cpx<R> *)out,
                                               (const cpx<R> *)tw, weights, a0, g,
                                               ldt);
  }
  return hipGetLastError();
}

// stage 0: k_adj_row (n = lNx), 1: k_adj_col (n = lNy), 2: k_adj_inv_col
// (n = lNy, aux = P), 3: k_adj_inv_row (n = lNx).
template <typename R>
hipError_t launch_adj_fft(int stage, int n, const void *in, void *out,
                          const void *tw, const void *aux, const double *prop_a,
                          const double *weights, int a0, const Geom &g, int na,
                          int ldt, hipStream_t st) {
  switch (n) {
    case 64: return adj_fft_launch<R, 64>(stage, in, out, tw, aux, prop_a, weights, a0, g, na, ldt, st);
    case 128: return adj_fft_launch<R, 128>(stage, in, out, tw, aux, prop_a, weights, a0, g, na, ldt, st);
    case 256: return adj_fft_launch<R, 256>(stage, in, out, tw, aux, prop_a, weights, a0, g, na, ldt, st);
    case 512: return adj_fft_launch<R, 512>(stage, in, out, tw, aux, prop_a, weights, a0, g, na, ldt, st);
  }
  return hipErrorInvalidValue;
}

#define ODT_ADJ_INST(R)                                                         \
  template hipError_t launch_adj_rotate<R>(int, const cpx<R> *, int,            \
                                           const double *, cpx<R> *, int, int,  \
                                           int, hipStream_t);                   \
  template hipError_t launch_scale_ends<R>(R *, long long, int, long long,      \
                                           long long, long long, long long,     \
                                           double, hipStream_t);                \
  template hipError_t launch_adj_fft<R>(int, int, const void *, void *,         \
                                        const void *, const void *,             \
                                        const double *, const double *, int,    \
                                        const Geom &, int, int, hipStream_t);

ODT_ADJ_INST(float)
ODT_ADJ_INST(double)

}  // namespace odt

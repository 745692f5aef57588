// gfx950 kernels of the filtered-backpropagation path.
//
// Data flow for one chunk of projections (reference: the angle loop of
// odtbrain/_alg3d_bpp.py:492-549, _alg3d_bppt.py:865-947, _alg2d_bpp.py:301-324):
//
//   k_fwd_rows   pad (np.pad edge / linear_ramp, weights folded in) and
//                forward transform along x; keeps only the x-frequencies
//                inside the low-pass support, stored column-major (F1)
//   k_fwd_cols   forward transform along y, multiply by the prefactor table
//                (ramp, low-pass, lD phase, 1/(lNx*lNy), and the fused
//                B-spline prefilter factors); keeps the support box (S)
//   k_col_ifft   per (x-frequency, z): propagator exp(i*km*(M-1)*z) (with
//                the fused z-prefilter), inverse transform along y, crop (T)
//   k_row_ifft   inverse transform along x, crop, fused x-prefilter; writes
//                the B-spline coefficient slab of the projection
//   k_rotate_lds2 / k_rotate_lds / k_affine_acc
//                resample every slab of the batch at the rotated position
//                of each output voxel and sum over angles (compensated)
//
// k_iir is the explicit separable prefilter (used when the prefilter is not
// folded into the transforms), k_fft_batch backs the standalone transform
// entry point.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "odt_device.h"
#include "odt_fft8.h"
#include "odt_launch.h"

namespace odt {

extern __shared__ __align__(16) unsigned char odt_smem[];

// ---------------------------------------------------------------------------
// twiddles: tw[m] = exp(+2 pi i m / n)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_twiddles(cpx<R> *tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[m] = mk<R>((R)c, (R)s);
}

// ---------------------------------------------------------------------------
// padding (numpy.pad semantics as called at _alg3d_bpp.py:496-506):
// axis 0 (y) is padded first within the original columns, then axis 1 (x)
// over all rows; edge mode replicates the nearest sample; linear_ramp goes
// from `end` at the outer border to the edge value, excluding the edge:
// value_i = end + i*(edge - end)/width, i = 0 at the outer border.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cpx<R> ramp_val(cpx<R> edge, cpx<R> end, int i,
                                           int width) {
  const R f = (R)i / (R)width;
  return mk<R>(end.re + f * (edge.re - end.re), end.im + f * (edge.im - end.im));
}

// Sample (j, col) of the caller's projection, widened to the compute type.
template <typename C, typename RI>
__device__ __forceinline__ cpx<C> ld_c(const cpx<RI> *u, int idx) {
  const cpx<RI> v = u[idx];
  return mk<C>((C)v.re, (C)v.im);
}

template <typename C, typename RI>
__device__ __forceinline__ cpx<C> ramp_row(const cpx<RI> *u, const Geom &g,
                                           C w, cpx<C> end, int j, int col) {
  const int Ny = g.Ny, Nx = g.Nx;
  if (j < g.padyl) {
    const cpx<C> e = w * ld_c<C>(u, col);
    return ramp_val(e, end, j, g.padyl);
  }
  if (j >= g.padyl + Ny) {
    const int padyr = g.lNy - Ny - g.padyl;
    const int r = j - g.padyl - Ny;
    const cpx<C> e = w * ld_c<C>(u, (Ny - 1) * Nx + col);
    return ramp_val(e, end, padyr - 1 - r, padyr);
  }
  return w * ld_c<C>(u, (j - g.padyl) * Nx + col);
}

template <typename C, typename RI>
__device__ __forceinline__ cpx<C> padded_value(const cpx<RI> *u, const Geom &g,
                                               C w, cpx<C> end, int j, int i) {
  const int Ny = g.Ny, Nx = g.Nx;
  if (g.pad_mode == 0) return w * ld_c<C>(u, j * Nx + i);
  if (g.pad_mode == 1) {
    int jj = j - g.padyl, ii = i - g.padxl;
    jj = jj < 0 ? 0 : (jj >= Ny ? Ny - 1 : jj);
    ii = ii < 0 ? 0 : (ii >= Nx ? Nx - 1 : ii);
    return w * ld_c<C>(u, jj * Nx + ii);
  }
  if (i < g.padxl) {
    const cpx<C> e = ramp_row<C>(u, g, w, end, j, 0);
    return ramp_val(e, end, i, g.padxl);
  }
  if (i >= g.padxl + Nx) {
    const int padxr = g.lNx - Nx - g.padxl;
    const int r = i - g.padxl - Nx;
    const cpx<C> e = ramp_row<C>(u, g, w, end, j, Nx - 1);
    return ramp_val(e, end, padxr - 1 - r, padxr);
  }
  return ramp_row<C>(u, g, w, end, j, i - g.padxl);
}

// ---------------------------------------------------------------------------
// forward transforms.  RI is the real type of the caller's projections and
// of the support spectrum S; C the type the pad, both transforms and the
// prefactor are computed in (float64 for both precisions: S is rounded to
// RI once, which keeps the forward stage's share of the float32 path's
// error at that one rounding).
// ---------------------------------------------------------------------------
template <typename RI, typename C>
__global__ void __launch_bounds__(256)
    k_fwd_rows(const cpx<RI> *__restrict__ usin,
               const double *__restrict__ weights, cpx<C> *__restrict__ F1,
               const cpx<C> *__restrict__ twx, Geom g, double pad_re,
               double pad_im, int a0, int nb) {
  cpx<C> *buf = reinterpret_cast<cpx<C> *>(odt_smem);
  cpx<C> *tmp = buf + nb * lds_len(g.lNx);
  const int ai = blockIdx.y;
  const int a = a0 + ai;
  // edge padding: the padded rows outside the detector repeat its first or
  // last row, so only the detector rows are transformed (the column stage
  // reads the others' transforms from the nearest detector row)
  const bool edge = g.pad_mode == 1;
  const int jend = edge ? g.padyl + g.Ny : g.lNy;
  const int j0 = blockIdx.x * nb + (edge ? g.padyl : 0);
  const cpx<RI> *u = usin + (size_t)a * g.Ny * g.Nx;
  const C w = (C)weights[a];
  const cpx<C> end = mk<C>((C)pad_re, (C)pad_im);
  for (int idx = threadIdx.x; idx < nb * g.lNx; idx += blockDim.x) {
    const int r = idx / g.lNx, i = idx - r * g.lNx;
    const int j = j0 + r;
    buf[idx] = j < jend ? padded_value<C>(u, g, w, end, j, i) : mk<C>(0, 0);
  }
  __syncthreads();
  const cpx<C> *res = lds_transform<C>(buf, tmp, g.lNx, g.logNx, nb, twx, -1);
  for (int idx = threadIdx.x; idx < g.Lx * nb; idx += blockDim.x) {
    const int c = idx / nb, r = idx - c * nb;
    const int j = j0 + r;
    if (j < jend) {
      const int bin = wrap(g.jx0 + c, g.lNx);
      F1[((size_t)ai * g.Lx + c) * g.lNy + j] = res[r * g.lNx + bin];
    }
  }
}

template <typename RI, typename C>
__global__ void __launch_bounds__(256)
    k_fwd_cols(const cpx<C> *__restrict__ F1, cpx<RI> *__restrict__ S,
               const cpx<C> *__restrict__ twy, const cpx<C> *__restrict__ P,
               Geom g, int nb) {
  cpx<C> *buf = reinterpret_cast<cpx<C> *>(odt_smem);
  cpx<C> *tmp = buf + nb * lds_len(g.lNy);
  const int ai = blockIdx.y;
  const int c0 = blockIdx.x * nb;
  const bool edge = g.pad_mode == 1;
  for (int idx = threadIdx.x; idx < nb * g.lNy; idx += blockDim.x) {
    const int r = idx / g.lNy, j = idx - r * g.lNy;
    const int c = c0 + r;
    const int jr = edge ? min(max(j, g.padyl), g.padyl + g.Ny - 1) : j;
    buf[idx] = c < g.Lx ? F1[((size_t)ai * g.Lx + c) * g.lNy + jr] : mk<C>(0, 0);
  }
  __syncthreads();
  const cpx<C> *res = lds_transform<C>(buf, tmp, g.lNy, g.logNy, nb, twy, -1);
  for (int idx = threadIdx.x; idx < nb * g.Ly; idx += blockDim.x) {
    const int r = idx / g.Ly, cy = idx - r * g.Ly;
    const int c = c0 + r;
    if (c < g.Lx) {
      const int bin = wrap(g.jy0 + cy, g.lNy);
      const cpx<C> v = res[r * g.lNy + bin] * P[(size_t)c * g.Ly + cy];
      S[((size_t)ai * g.Lx + c) * g.Ly + cy] = mk<RI>((RI)v.re, (RI)v.im);
    }
  }
}

// Propagator table g_k(cx, cy) of every slice k and support point: the
// propagator exp(i a z_k), a = km (M - 1) from the float64 table, evaluated
// with float64 sin/cos, times the fused z-prefilter terms where present
// (see k_col_ifft), rounded once.  The same for every angle, so the column
// stage reads it instead of evaluating it per angle.
template <typename R>
__global__ void __launch_bounds__(256)
    k_prop_table(const double *__restrict__ prop_a, const PropAux<R> *__restrict__ aux,
                 const R *__restrict__ q0z, Geom g, cpx<R> *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)g.Lx * g.Ly;
  if (i >= per * g.ln) return;
  const int k = (int)(i / per);
  const long long t = i - (long long)k * per;
  double sn, cs;
  sincos(prop_a[t] * (g.z0 + (double)k), &sn, &cs);
  double re = cs, im = sn;
  if (g.fuse_z) {
    const PropAux<R> x = aux[t];
    const double ib = (double)x.ib, qa = (double)q0z[k], qb = (double)q0z[g.ln - 1 - k];
    re = ib * cs + qa * (double)x.r0.re + qb * (double)x.r1.re;
    im = ib * sn + qa * (double)x.r0.im + qb * (double)x.r1.im;
  }
  out[i] = mk<R>((R)re, (R)im);
}

template <typename R>
hipError_t launch_prop_table(const double *pa, const PropAux<R> *aux, const R *q0z,
                             const Geom &g, cpx<R> *out, hipStream_t st) {
  const long long n = (long long)g.Lx * g.Ly * g.ln;
  k_prop_table<R><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(pa, aux, q0z, g, out);
  return hipGetLastError();
}

// Column stage: for x-frequency column c and slices k0..k0+nb-1 build
// S*g_k along the y-frequencies (zero outside the support), inverse
// transform along y, crop rows padyl..padyl+Ny-1 (plus the fused
// y-prefilter boundary terms for the tilted path), write T[a][k][y][c].
template <typename R>
__global__ void __launch_bounds__(256)
    k_col_ifft(const cpx<R> *__restrict__ S, cpx<R> *__restrict__ T,
               const cpx<R> *__restrict__ twy,
               const double *__restrict__ prop_a,
               const PropAux<R> *__restrict__ aux,
               const R *__restrict__ q0z, const R *__restrict__ q0y, Geom g,
               int nb, int ldt, const cpx<R> *__restrict__ gtab) {
  cpx<R> *buf = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *tmp = buf + nb * lds_len(g.lNy);
  __shared__ cpx<R> corr[2 * 16];
  const int c = blockIdx.x;
  const int k0 = blockIdx.y * nb;
  const int ai = blockIdx.z;
  const int ln = g.ln;
  for (int idx = threadIdx.x; idx < nb * g.lNy; idx += blockDim.x)
    buf[idx] = mk<R>(0, 0);
  __syncthreads();
  const cpx<R> *Sc = S + ((size_t)ai * g.Lx + c) * g.Ly;
  for (int idx = threadIdx.x; idx < nb * g.Ly; idx += blockDim.x) {
    const int r = idx / g.Ly, cy = idx - r * g.Ly;
    const int k = k0 + r;
    if (k >= ln) continue;
    const size_t t = (size_t)c * g.Ly + cy;
    cpx<R> gk;
    if (gtab) {
      gk = gtab[(size_t)k * g.Lx * g.Ly + t];
    } else {
      const cpx<R> e = phasor(prop_a[t] * (g.z0 + (double)k), R(0));
      gk = e;
      if (g.fuse_z) {
        const PropAux<R> x = aux[t];
        gk = x.ib * e + q0z[k] * x.r0 + q0z[ln - 1 - k] * x.r1;
      }
    }
    buf[r * g.lNy + wrap(g.jy0 + cy, g.lNy)] = Sc[cy] * gk;
  }
  __syncthreads();
  const cpx<R> *res = lds_transform<R>(buf, tmp, g.lNy, g.logNy, nb, twy, +1);
  const int ny = g.Ny;
  if (g.fuse_y) {
    if (threadIdx.x < nb) {
      const cpx<R> *v = res + threadIdx.x * g.lNy;
      const R b1 = (R)g.b1;
      corr[2 * threadIdx.x] =
          b1 * (v[wrap(g.padyl - 1, g.lNy)] - v[wrap(g.padyl + 1, g.lNy)]);
      corr[2 * threadIdx.x + 1] = b1 * (v[wrap(g.padyl + ny, g.lNy)] -
                                        v[wrap(g.padyl + ny - 2, g.lNy)]);
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < nb * ny; idx += blockDim.x) {
    const int r = idx / ny, y = idx - r * ny;
    const int k = k0 + r;
    if (k >= ln) continue;
    cpx<R> v = res[r * g.lNy + wrap(g.padyl + y, g.lNy)];
    if (g.fuse_y)
      v = v + q0y[y] * corr[2 * r] + q0y[ny - 1 - y] * corr[2 * r + 1];
    T[(((size_t)ai * ln + k) * ny + y) * ldt + c] = v;
  }
}

// Non-temporal stores for the streams written once and read back by a later
// kernel from HBM (T, the coefficient slabs): they keep the write stream
// from evicting the L2 lines that the writing kernel still reads (measured
// at 256^2 x 360: 58.1 -> 54.6 ms per reconstruction).
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(cpx<float> *p, cpx<float> v) {
  f2v x = {v.re, v.im}; __builtin_nontemporal_store(x, reinterpret_cast<f2v *>(p));
}
// 16-byte non-temporal stores measured slow (256^2 x 360 in FP64: 155 ->
// 346 ms per reconstruction): complex128 samples take plain stores.
__device__ __forceinline__ void st_nt(cpx<double> *p, cpx<double> v) { *p = v; }
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }

// Column stage of the 2D path (lNy = Ny = 1, no y-prefilter): the
// transform along y is the identity, so T[a][k][0][c] = S[a][c] * g_k(c)
// with g_k as in k_col_ifft.  Threads take consecutive columns (coalesced
// rows of T), blocks runs of kb slices.
template <typename R>
__global__ void __launch_bounds__(256)
    k_prop_2d(const cpx<R> *__restrict__ S, cpx<R> *__restrict__ T,
              const double *__restrict__ prop_a,
              const PropAux<R> *__restrict__ aux, const R *__restrict__ q0z,
              Geom g, int kb, int ldt, const cpx<R> *__restrict__ gtab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int k0 = blockIdx.y * kb, ai = blockIdx.z;
  if (c >= g.Lx) return;
  const int ln = g.ln;
  const cpx<R> s = S[(size_t)ai * g.Lx + c];
  const double pa = prop_a[c];
  PropAux<R> x{};
  if (g.fuse_z) x = aux[c];
  const int kend = min(k0 + kb, ln);
  for (int k = k0; k < kend; ++k) {
    cpx<R> gk;
    if (gtab) {
      gk = gtab[(size_t)k * g.Lx + c];
    } else {
      const cpx<R> e = phasor(pa * (g.z0 + (double)k), R(0));
      gk = e;
      if (g.fuse_z) gk = x.ib * e + q0z[k] * x.r0 + q0z[ln - 1 - k] * x.r1;
    }
    st_nt(T + ((size_t)ai * ln + k) * ldt + c, s * gk);
  }
}

template <typename R>
__device__ __forceinline__ void store_val(R *p, cpx<R> v) { st_nt(p, v.re); }
template <typename R>
__device__ __forceinline__ void store_val(cpx<R> *p, cpx<R> v) { st_nt(p, v); }

// Signed frequency index of FFT bin `bin` (numpy.fft.fftfreq order).
template <int N>
__device__ __forceinline__ int freq_of_bin(int bin) {
  return bin < (N + 1) / 2 ? bin : bin - N;
}

// Fused-prefilter boundary terms of one transform: the owners of the four
// outputs Y[p-1], Y[p+1], Y[p+n-2], Y[p+n] (indices mod N, p = crop start,
// n = crop length) publish them in `corr`; returns
// R0 = b1 (Y[p-1] - Y[p+1]) and R1 = b1 (Y[p+n] - Y[p+n-2]).
// WAVE: the transform lives in one wavefront (no block barrier needed).
template <typename R, int N, int TT, bool WAVE>
__device__ __forceinline__ void boundary_terms(const cpx<R> *v, int j, int p,
                                               int n, R b1, cpx<R> *corr,
                                               cpx<R> &R0, cpx<R> &R1) {
  const int sb[4] = {(p - 1 + N) & (N - 1), (p + 1) & (N - 1),
                     (p + n - 2 + N) & (N - 1), (p + n) & (N - 1)};
  if (!WAVE) __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int bin = j + m * TT;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (bin == sb[i]) corr[i] = v[m];
  }
  if (WAVE)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  else
    __syncthreads();
  R0 = b1 * (corr[0] - corr[1]);
  R1 = b1 * (corr[3] - corr[2]);
  if (WAVE) __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ float rdlane(float x, int l) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
__device__ __forceinline__ double rdlane(double x, int l) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// Value of lane `src` (an absolute lane of the wavefront) for every lane.
__device__ __forceinline__ float permlane(float x, int src) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double permlane(double x, int src) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffff));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// Lane l of this transform's lane group (TT lanes, TT divides 64).
template <int TT, typename R>
__device__ __forceinline__ R group_lane(R x, int l) {
  if constexpr (TT == 64) return rdlane(x, l);
  else return permlane(x, (int)(threadIdx.x & (64 - TT)) + l);
}
// Output bin `bin` (the same for every transform of the wavefront) of a
// transform held in the strided layout v[m] = X[j + m*TT] by TT lanes:
// slot bin/TT of lane bin%TT of the group.
template <typename R, int TT>
__device__ __forceinline__ cpx<R> wave_bin(const cpx<R> *v, int bin) {
  const int m = __builtin_amdgcn_readfirstlane(bin / TT);
  const int l = bin - m * TT;
  // uniform branches rather than x = v[m], which the compiler would turn
  // into a dynamically indexed (scratch) array
  switch (m) {
    case 0: return mk<R>(group_lane<TT>(v[0].re, l), group_lane<TT>(v[0].im, l));
    case 1: return mk<R>(group_lane<TT>(v[1].re, l), group_lane<TT>(v[1].im, l));
    case 2: return mk<R>(group_lane<TT>(v[2].re, l), group_lane<TT>(v[2].im, l));
    case 3: return mk<R>(group_lane<TT>(v[3].re, l), group_lane<TT>(v[3].im, l));
    case 4: return mk<R>(group_lane<TT>(v[4].re, l), group_lane<TT>(v[4].im, l));
    case 5: return mk<R>(group_lane<TT>(v[5].re, l), group_lane<TT>(v[5].im, l));
    case 6: return mk<R>(group_lane<TT>(v[6].re, l), group_lane<TT>(v[6].im, l));
    default: return mk<R>(group_lane<TT>(v[7].re, l), group_lane<TT>(v[7].im, l));
  }
}
// Fused-prefilter boundary terms of a transform held by one wavefront (see
// boundary_terms), read across lanes without LDS.
template <typename R, int N, int TT>
__device__ __forceinline__ void wave_boundary_terms(const cpx<R> *v, int p,
                                                    int n, R b1, cpx<R> &R0,
                                                    cpx<R> &R1) {
  const cpx<R> y0 = wave_bin<R, TT>(v, (p - 1 + N) & (N - 1));
  const cpx<R> y1 = wave_bin<R, TT>(v, (p + 1) & (N - 1));
  const cpx<R> y2 = wave_bin<R, TT>(v, (p + n - 2 + N) & (N - 1));
  const cpx<R> y3 = wave_bin<R, TT>(v, (p + n) & (N - 1));
  R0 = b1 * (y0 - y1);
  R1 = b1 * (y3 - y2);
}

// Column stage, register radix-8 engine (power-of-two lNy = N).
// A block owns CPB = 16 adjacent support columns c0..c0+15 and the slices
// k0..k0+kb-1; its tpb transform slots take the columns tpb at a time.  The
// S column and the tables are read from cache per use (keeping them out of
// the transform's registers).  Outputs are transposed through an LDS tile so
// that each row of T[a][k][y][c] (stride ldt) is written in contiguous
// 16-column (128-byte) segments; without room for the tile (very long
// columns) the transform writes T directly.
// BAND: the y support lies inside (-N/4, N/4), so of the eight register
// slots (bins j + m*N/8) only m = 0, 1, 6, 7 hold inputs; the first radix-8
// stage skips the others (half the loads and propagators).
// Loads are predicated with selects on clamped indices, and when the crop
// window is aligned to N/8 the crop test is uniform across the wavefront.
// Forward stage with the register radix-8 engine (power-of-two lengths 64
// to 2048; longer ones and other lengths keep k_fwd_rows / k_fwd_cols):
// one padded row (or one support column) per transform, 256 / (N/8)
// transforms per block.  Rows: the transform's outputs go through its LDS
// buffer so that the block writes F1's support columns with its rows
// consecutive, as k_fwd_rows does.  Columns: the outputs of the support
// bins times the prefactor go straight to S.
template <typename RI, typename C, int N>
__global__ void __launch_bounds__(256)
    k_fwd_rows8(const cpx<RI> *__restrict__ usin, const double *__restrict__ weights,
                cpx<C> *__restrict__ F1, const cpx<C> *__restrict__ twx, Geom g,
                double pad_re, double pad_im, int a0) {
  using FFT = Fft8<C, N, -1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 256, "one block holds a whole transform");
  constexpr int TPB = 256 / TT;
  const int t = threadIdx.x / TT, jj = threadIdx.x - t * TT;
  cpx<C> *lds = reinterpret_cast<cpx<C> *>(odt_smem);
  cpx<C> *b0 = lds + FFT::NBUF * t * FFT::PADN, *b1 = b0 + FFT::PADN;
  const int ai = blockIdx.y, a = a0 + ai;
  const bool edge = g.pad_mode == 1;
  const int jend = edge ? g.padyl + g.Ny : g.lNy;
  const int jb = blockIdx.x * TPB + (edge ? g.padyl : 0);
  const int j = jb + t;
  const cpx<RI> *u = usin + (size_t)a * g.Ny * g.Nx;
  const C w = (C)weights[a];
  const cpx<C> end = mk<C>((C)pad_re, (C)pad_im);
  cpx<C> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
    v[m] = j < jend ? padded_value<C>(u, g, w, end, j, jj + m * TT) : mk<C>(0, 0);
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twx, jj);
  FFT::template run<false>(v, b0, b1, tw, jj);
  __syncthreads();  // every transform's last stage has read its buffers
#pragma unroll
  for (int m = 0; m < 8; ++m) b0[lds_pad(jj + m * TT)] = v[m];
  __syncthreads();
  for (int idx = threadIdx.x; idx < g.Lx * TPB; idx += 256) {
    const int c = idx / TPB, r = idx - c * TPB;
    if (jb + r < jend) {
      const cpx<C> *res = lds + FFT::NBUF * r * FFT::PADN;
      F1[((size_t)ai * g.Lx + c) * g.lNy + jb + r] = res[lds_pad(wrap(g.jx0 + c, N))];
    }
  }
}

template <typename RI, typename C, int N>
__global__ void __launch_bounds__(256)
    k_fwd_cols8(const cpx<C> *__restrict__ F1, cpx<RI> *__restrict__ S,
                const cpx<C> *__restrict__ twy, const cpx<C> *__restrict__ P, Geom g) {
  using FFT = Fft8<C, N, -1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 256, "one block holds a whole transform");
  constexpr int TPB = 256 / TT;
  const int t = threadIdx.x / TT, jj = threadIdx.x - t * TT;
  cpx<C> *lds = reinterpret_cast<cpx<C> *>(odt_smem);
  cpx<C> *b0 = lds + FFT::NBUF * t * FFT::PADN, *b1 = b0 + FFT::PADN;
  const int ai = blockIdx.y;
  const int c = blockIdx.x * TPB + t;
  const bool live = c < g.Lx;
  const bool edge = g.pad_mode == 1;
  const cpx<C> *col = F1 + ((size_t)ai * g.Lx + (live ? c : 0)) * g.lNy;
  cpx<C> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int jr0 = jj + m * TT;
    // edge padding: the padded rows hold the nearest detector row
    const int jr = edge ? min(max(jr0, g.padyl), g.padyl + g.Ny - 1) : jr0;
    v[m] = live ? col[jr] : mk<C>(0, 0);
  }
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twy, jj);
  FFT::template run<false>(v, b0, b1, tw, jj);
  if (!live) return;  // (after the transform's barriers)
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cy = freq_of_bin<N>(jj + m * TT) - g.jy0;
    if (cy >= 0 && cy < g.Ly) {
      const cpx<C> x = v[m] * P[(size_t)c * g.Ly + cy];
      S[((size_t)ai * g.Lx + c) * g.Ly + cy] = mk<RI>((RI)x.re, (RI)x.im);
    }
  }
}

constexpr int kColCpb = 16;

template <int N, bool BAND>
__device__ __forceinline__ constexpr bool slot_used(int m) {
  return !BAND || m < 2 || m > 5;
}

template <typename R, int N, bool BAND>
__global__ void __launch_bounds__(512)
    k_col_fft8(const cpx<R> *__restrict__ S, cpx<R> *__restrict__ T,
               const cpx<R> *__restrict__ twy,
               const double *__restrict__ prop_a,
               const PropAux<R> *__restrict__ aux,
               const R *__restrict__ q0z, const R *__restrict__ q0y, Geom g,
               int kb, int ldt, int staged, const cpx<R> *__restrict__ gtab) {
  using FFT = Fft8<R, N, +1>;
  constexpr int TT = FFT::T;
  constexpr bool WAVE = TT <= 64;
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *lds = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *b0 = lds + FFT::NBUF * t * FFT::PADN, *b1 = b0 + FFT::PADN;
  cpx<R> *corr = lds + FFT::NBUF * tpb * FFT::PADN + 4 * t;
  cpx<R> *tile = lds + FFT::NBUF * tpb * FFT::PADN + 4 * tpb;  // ny x (CPB+1)
  const int c0 = blockIdx.x * kColCpb;
  const int k0 = blockIdx.y * kb, ai = blockIdx.z;
  const int ln = g.ln, ny = g.Ny;
  const int kend = min(k0 + kb, ln);
  // per-slot support index (clamped) and its validity: loop invariant
  int cyc[8];
  bool cyv[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cy = freq_of_bin<N>(j + m * TT) - g.jy0;
    cyv[m] = cy >= 0 && cy < g.Ly;
    cyc[m] = cyv[m] ? cy : 0;
  }
  const bool uni = g.padyl % TT == 0 && ny % TT == 0;
  const int mlo = g.padyl / TT, mhi = (g.padyl + ny) / TT;
  // the fused y-prefilter's boundary weights of the lane's output rows (the
  // same for every transform)
  R wya[8], wyb[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int y = j + m * TT - g.padyl;
    const bool in = g.fuse_y && (uni ? (m >= mlo && m < mhi) : (y >= 0 && y < ny));
    wya[m] = in ? q0y[y] : R(0);
    wyb[m] = in ? q0y[ny - 1 - y] : R(0);
  }
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twy, j);
  for (int k = k0; k < kend; ++k) {
    const double z = g.z0 + (double)k;
    // z-prefilter boundary terms only where q0 is not negligible
    const R qa = g.fuse_z ? q0z[k] : R(0);
    const R qb = g.fuse_z ? q0z[ln - 1 - k] : R(0);
    const bool edge = qa != R(0) || qb != R(0);
    cpx<R> *trow = T + (((size_t)ai * ln + k) * ny) * ldt;
    for (int cb = 0; cb < kColCpb; cb += tpb) {
      const int cl = cb + t;  // local column of this slot
      const int c = c0 + cl;
      const bool live = cl < kColCpb && c < g.Lx;
      const int cs = live ? c : 0;
      const cpx<R> *Sc = S + ((size_t)ai * g.Lx + cs) * g.Ly;
      const size_t tr = (size_t)cs * g.Ly;
      cpx<R> v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (!slot_used<N, BAND>(m)) continue;
        const int cy = cyc[m];
        cpx<R> gk;
        if (gtab) {
          gk = gtab[(size_t)k * g.Lx * g.Ly + tr + cy];
        } else {
          gk = phasor(prop_a[tr + cy] * z, R(0));
          if (g.fuse_z) {
            const PropAux<R> *x = aux + tr + cy;
            gk = x->ib * gk;
            if (edge) gk = gk + qa * x->r0 + qb * x->r1;
          }
        }
        const cpx<R> val = Sc[cy] * gk;
        const bool ok = live && cyv[m];
        v[m] = mk<R>(ok ? val.re : R(0), ok ? val.im : R(0));
      }
      FFT::template run<BAND>(v, b0, b1, tw, j);
      cpx<R> R0 = mk<R>(0, 0), R1 = mk<R>(0, 0);
      if (g.fuse_y) {
        if constexpr (WAVE)
          wave_boundary_terms<R, N, TT>(v, g.padyl, ny, (R)g.b1, R0, R1);
        else
          boundary_terms<R, N, TT, WAVE>(v, j, g.padyl, ny, (R)g.b1, corr, R0,
                                         R1);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int y = j + m * TT - g.padyl;
        const bool in = uni ? (m >= mlo && m < mhi) : (y >= 0 && y < ny);
        if (live && in) {
          cpx<R> val = v[m];
          if (g.fuse_y) val = val + wya[m] * R0 + wyb[m] * R1;
          if (staged)
            tile[y * (kColCpb + 1) + cl] = val;
          else
            st_nt(trow + (size_t)y * ldt + c, val);
        }
      }
    }
    if (staged) {
      __syncthreads();
      for (int idx = threadIdx.x; idx < ny * kColCpb; idx += blockDim.x) {
        const int y = idx / kColCpb, cl = idx - y * kColCpb;
        if (c0 + cl < g.Lx)
          st_nt(trow + (size_t)y * ldt + c0 + cl, tile[y * (kColCpb + 1) + cl]);
      }
      __syncthreads();
    }
  }
}

// Column stage with the propagator held in registers across the angles of
// a chunk (power-of-two lNy = N <= 512, band-limited y support, propagator
// table present).  The propagator g_k(c, cy) of the per-plan table is the
// same for every angle, so a block owns 16 support columns of one slice k
// for all na angles of the launch: it loads its g values once and then, per
// angle, reads the S columns (shared by every block of that angle, L2
// resident), transforms and writes its T rows through the LDS tile.  The
// table is read once per launch instead of once per angle.  Each transform
// slot t of the block takes the columns t, t + TPB, ...; the next
// transform's S values are loaded while the current one runs.
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_col_fft8_ang(const cpx<R> *__restrict__ S, cpx<R> *__restrict__ T,
                   const cpx<R> *__restrict__ twy, const R *__restrict__ q0y,
                   Geom g, int na, int ldt, const cpx<R> *__restrict__ gtab) {
  using FFT = Fft8<R, N, +1>;
  constexpr int TT = FFT::T;
  static_assert(TT <= 64, "one wavefront per transform");
  constexpr int TPB = 256 / TT < kColCpb ? 256 / TT : kColCpb;
  constexpr int NT = kColCpb / TPB;  // transforms per slot and angle
  constexpr int MU[4] = {0, 1, 6, 7};  // band slots holding inputs
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *lds = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *b0 = lds + t * FFT::PADN;
  cpx<R> *tile = lds + TPB * FFT::PADN;  // ny x (CPB+1)
  const int c0 = blockIdx.x * kColCpb, k = blockIdx.y;
  const int ln = g.ln, ny = g.Ny;
  int cyc[4];
  bool cyv[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int cy = freq_of_bin<N>(j + MU[u] * TT) - g.jy0;
    cyv[u] = cy >= 0 && cy < g.Ly;
    cyc[u] = cyv[u] ? cy : 0;
  }
  const bool uni = g.padyl % TT == 0 && ny % TT == 0;
  const int mlo = g.padyl / TT, mhi = (g.padyl + ny) / TT;
  // the fused y-prefilter's boundary weights of the lane's output rows (the
  // same for every transform)
  R wya[8], wyb[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int y = j + m * TT - g.padyl;
    const bool in = g.fuse_y && (uni ? (m >= mlo && m < mhi) : (y >= 0 && y < ny));
    wya[m] = in ? q0y[y] : R(0);
    wyb[m] = in ? q0y[ny - 1 - y] : R(0);
  }
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twy, j);
  // this slot's propagator values, all angles
  cpx<R> gk[NT][4];
  int ccl[NT];
  bool live[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const int cl = t + i * TPB;
    const int c = c0 + cl;
    live[i] = c < g.Lx;
    ccl[i] = live[i] ? c : 0;
    const cpx<R> *gc = gtab + (size_t)k * g.Lx * g.Ly + (size_t)ccl[i] * g.Ly;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const cpx<R> v = gc[cyc[u]];
      const bool ok = live[i] && cyv[u];
      gk[i][u] = mk<R>(ok ? v.re : R(0), ok ? v.im : R(0));
    }
  }
  auto load_s = [&](int ai, int i, cpx<R> *dst) {
    const cpx<R> *Sc = S + ((size_t)ai * g.Lx + ccl[i]) * g.Ly;
#pragma unroll
    for (int u = 0; u < 4; ++u) dst[u] = Sc[cyc[u]];
  };
  cpx<R> sn[4];
  load_s(0, 0, sn);
  for (int ai = 0; ai < na; ++ai) {
    cpx<R> *trow = T + (((size_t)ai * ln + k) * ny) * ldt;
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      cpx<R> sv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) sv[u] = sn[u];
      // prefetch the next transform's S values (clamped at the end)
      const int ni = i + 1 < NT ? i + 1 : 0;
      const int na2 = i + 1 < NT ? ai : (ai + 1 < na ? ai + 1 : ai);
      load_s(na2, ni, sn);
      cpx<R> v[8];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const cpx<R> val = sv[u] * gk[i][u];
        const bool ok = live[i] && cyv[u];
        v[MU[u]] = mk<R>(ok ? val.re : R(0), ok ? val.im : R(0));
      }
      FFT::template run<true>(v, b0, b0, tw, j);
      cpx<R> R0 = mk<R>(0, 0), R1 = mk<R>(0, 0);
      if (g.fuse_y) wave_boundary_terms<R, N, TT>(v, g.padyl, ny, (R)g.b1, R0, R1);
      const int cl = t + i * TPB;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int y = j + m * TT - g.padyl;
        const bool in = uni ? (m >= mlo && m < mhi) : (y >= 0 && y < ny);
        if (live[i] && in) {
          cpx<R> val = v[m];
          if (g.fuse_y) val = val + wya[m] * R0 + wyb[m] * R1;
          tile[y * (kColCpb + 1) + cl] = val;
        }
      }
    }
    __syncthreads();
    {
      // the tile out in 128-byte row segments: a thread keeps its column
      // and takes every RPP-th row, four rows' LDS reads before their stores
      constexpr int NTH = TPB * TT;  // the block's threads
      constexpr int RPP = NTH / kColCpb;
      static_assert(NTH % kColCpb == 0, "whole row segments per pass");
      const int cl = threadIdx.x % kColCpb;
      if (c0 + cl < g.Lx) {
        const cpx<R> *src = tile + cl;
        cpx<R> *dst = trow + c0 + cl;
        int y = threadIdx.x / kColCpb;
        for (; y + 3 * RPP < ny; y += 4 * RPP) {
          cpx<R> v4[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) v4[q] = src[(y + q * RPP) * (kColCpb + 1)];
#pragma unroll
          for (int q = 0; q < 4; ++q) st_nt(dst + (size_t)(y + q * RPP) * ldt, v4[q]);
        }
        for (; y < ny; y += RPP) st_nt(dst + (size_t)y * ldt, src[y * (kColCpb + 1)]);
      }
    }
    __syncthreads();
  }
}

// Row stage, register radix-8 engine (power-of-two lNx = N).  Each
// transform reads one row T[a][k][y][0..Lx) (contiguous, coalesced),
// transforms it, crops with the fused x-prefilter terms and writes one slab
// row.  Block: rows y0..y0+rb-1 of slice k, tpb rows at a time.  BAND and
// the load/crop predication as in the column stage.
//
// S2 non-null (2D, rows viewed as slices): there is no T; row r of angle a is
// S2[a][c] * gt[r][c], the support spectrum times the plan's propagator
// table (k_prop_2d's product, formed on load).
//
// float32 transforms of 1024 points and more (two LDS buffers, a barrier per
// stage) are held to 128 VGPRs, four waves per SIMD: two 512-thread blocks
// per CU instead of one, at the price of a few spilled loop invariants
// (measured on the batch of mixed 2D/3D reconstructions: 136.8 -> 133.7 ms
// per replay).  The same cap on the column stage (166 VGPRs) costs more in
// spills than it gains.
template <typename R, typename V, int N, bool BAND>
__global__ void __launch_bounds__(512)
    __attribute__((amdgpu_waves_per_eu(N >= 1024 && sizeof(R) == 4 ? 4 : 1)))
    k_row_fft8(const cpx<R> *__restrict__ T, V *__restrict__ slab,
               const cpx<R> *__restrict__ twx, const R *__restrict__ q0x,
               Geom g, int rb, int slab0, int ldt, const cpx<R> *__restrict__ S2,
               const cpx<R> *__restrict__ gt) {
  using FFT = Fft8<R, N, +1>;
  constexpr int TT = FFT::T;
  constexpr bool WAVE = TT <= 64;
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *lds = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *b0 = lds + FFT::NBUF * t * FFT::PADN, *b1 = b0 + FFT::PADN;
  cpx<R> *corr = lds + FFT::NBUF * tpb * FFT::PADN + 4 * t;
  const int y0 = blockIdx.x * rb, k = blockIdx.y, ai = blockIdx.z;
  const int ln = g.ln, ny = g.Ny, nx = g.Nx;
  int cxc[8];
  bool cxv[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int cx = freq_of_bin<N>(j + m * TT) - g.jx0;
    cxv[m] = cx >= 0 && cx < g.Lx;
    cxc[m] = cxv[m] ? cx : 0;
  }
  const bool uni = g.padxl % TT == 0 && nx % TT == 0;
  const int mlo = g.padxl / TT, mhi = (g.padxl + nx) / TT;
  typename FFT::Twid tw;
  FFT::load_twiddles(tw, twx, j);
  const int iters = (rb + tpb - 1) / tpb;
  // the fused x-prefilter's boundary weights of the lane's output columns
  // (the same for every row)
  R qa[8], qb[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int x = j + m * TT - g.padxl;
    const bool in = g.fuse_x && (uni ? (m >= mlo && m < mhi) : (x >= 0 && x < nx));
    qa[m] = in ? q0x[x] : R(0);
    qb[m] = in ? q0x[nx - 1 - x] : R(0);
  }
  // fused 2D rows: this angle's spectrum values of the lane's slots
  cpx<R> s2[8];
  if (S2) {
#pragma unroll
    for (int m = 0; m < 8; ++m)
      s2[m] = slot_used<N, BAND>(m) ? S2[(size_t)ai * g.Lx + cxc[m]] : mk<R>(0, 0);
  }
  // the next row's support values are loaded while this row is transformed
  cpx<R> vn[8];
  auto load_row = [&](int it2, cpx<R> *dst) {
    const int r2 = it2 * tpb + t;
    const int y2 = y0 + r2;
    const bool ok2 = it2 < iters && r2 < rb && y2 < ny;
    const cpx<R> *src = S2 ? gt + (size_t)(ok2 ? y2 : 0) * g.Lx
                           : T + (((size_t)ai * ln + k) * ny + (ok2 ? y2 : 0)) * ldt;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if (!slot_used<N, BAND>(m)) continue;
      cpx<R> val = src[cxc[m]];
      if (S2) val = s2[m] * val;
      const bool ok = ok2 && cxv[m];
      dst[m] = mk<R>(ok ? val.re : R(0), ok ? val.im : R(0));
    }
  };
  load_row(0, vn);
  for (int it = 0; it < iters; ++it) {
    const int r = it * tpb + t;
    const int y = y0 + r;
    const bool valid = r < rb && y < ny;
    if (WAVE && !valid) break;  // rows are taken in order: nothing left
    cpx<R> v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = vn[m];
    load_row(it + 1, vn);
    FFT::template run<BAND>(v, b0, b1, tw, j);
    cpx<R> R0 = mk<R>(0, 0), R1 = mk<R>(0, 0);
    if (g.fuse_x) {
      if constexpr (WAVE)
        wave_boundary_terms<R, N, TT>(v, g.padxl, nx, (R)g.b1, R0, R1);
      else
        boundary_terms<R, N, TT, WAVE>(v, j, g.padxl, nx, (R)g.b1, corr, R0,
                                       R1);
    }
    if (valid) {
      V *dst = slab + (((size_t)(slab0 + ai) * ny + y) * ln + k) * nx;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int x = j + m * TT - g.padxl;
        const bool in = uni ? (m >= mlo && m < mhi) : (x >= 0 && x < nx);
        if (in) {
          cpx<R> val = v[m];
          if (g.fuse_x) val = val + qa[m] * R0 + qb[m] * R1;
          store_val(dst + x, val);
        }
      }
    }
  }
}

// Row stage (generic transform length): rows y0..y0+nb-1 of slice k:
// inverse transform along x of the support columns of T, crop columns
// padxl..padxl+Nx-1 with the fused x-prefilter boundary terms, write the
// coefficient slab [s][y][k][x].
template <typename R, typename V>
__global__ void __launch_bounds__(256)
    k_row_ifft(const cpx<R> *__restrict__ T, V *__restrict__ slab,
               const cpx<R> *__restrict__ twx, const R *__restrict__ q0x,
               Geom g, int nb, int slab0, int ldt) {
  cpx<R> *buf = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *tmp = buf + nb * lds_len(g.lNx);
  __shared__ cpx<R> corr[2 * 16];
  const int y0 = blockIdx.x * nb;
  const int k = blockIdx.y;
  const int ai = blockIdx.z;
  const int ny = g.Ny, nx = g.Nx, ln = g.ln;
  for (int idx = threadIdx.x; idx < nb * g.lNx; idx += blockDim.x)
    buf[idx] = mk<R>(0, 0);
  __syncthreads();
  const cpx<R> *Tk = T + ((size_t)ai * ln + k) * ny * ldt;
  for (int idx = threadIdx.x; idx < g.Lx * nb; idx += blockDim.x) {
    const int r = idx / g.Lx, c = idx - r * g.Lx;
    const int y = y0 + r;
    if (y < ny) buf[r * g.lNx + wrap(g.jx0 + c, g.lNx)] = Tk[(size_t)y * ldt + c];
  }
  __syncthreads();
  const cpx<R> *res = lds_transform<R>(buf, tmp, g.lNx, g.logNx, nb, twx, +1);
  if (g.fuse_x) {
    if (threadIdx.x < nb) {
      const cpx<R> *v = res + threadIdx.x * g.lNx;
      const R b1 = (R)g.b1;
      corr[2 * threadIdx.x] =
          b1 * (v[wrap(g.padxl - 1, g.lNx)] - v[wrap(g.padxl + 1, g.lNx)]);
      corr[2 * threadIdx.x + 1] = b1 * (v[wrap(g.padxl + nx, g.lNx)] -
                                        v[wrap(g.padxl + nx - 2, g.lNx)]);
    }
    __syncthreads();
  }
  const int s = slab0 + ai;
  for (int idx = threadIdx.x; idx < nb * nx; idx += blockDim.x) {
    const int r = idx / nx, x = idx - r * nx;
    const int y = y0 + r;
    if (y >= ny) continue;
    cpx<R> v = res[r * g.lNx + wrap(g.padxl + x, g.lNx)];
    if (g.fuse_x)
      v = v + q0x[x] * corr[2 * r] + q0x[nx - 1 - x] * corr[2 * r + 1];
    store_val(slab + (((size_t)s * ny + y) * ln + k) * nx + x, v);
  }
}

// ---------------------------------------------------------------------------
// accumulation helpers of the resampling kernels (scipy.ndimage.rotate /
// affine_transform of every slab, summed over the angles of the batch,
// _alg3d_bpp.py:68-81 and 541-546, _alg3d_bppt.py:931-944)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void acc_zero(R &v) { v = R(0); }
template <typename R>
__device__ __forceinline__ void acc_zero(cpx<R> &v) { v = mk<R>(0, 0); }
// float64 running sums for the angle accumulation: wide_t<V> holds V in
// float64; the float32 compensation kept in HBM between launches is
// (rounded - exact), as in Kahan's scheme.
template <typename V> struct wide_of { using type = double; };
template <> struct wide_of<cpx<float>> { using type = cpx<double>; };
template <> struct wide_of<cpx<double>> { using type = cpx<double>; };
template <typename V> using wide_t = typename wide_of<V>::type;
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ cpx<double> widen(cpx<float> v) {
  return mk<double>((double)v.re, (double)v.im);
}
__device__ __forceinline__ cpx<double> widen(cpx<double> v) { return v; }
template <typename V>
__device__ __forceinline__ V narrow(double v) { return (V)v; }
template <typename V>
__device__ __forceinline__ V narrow(cpx<double> v);
template <>
__device__ __forceinline__ cpx<float> narrow<cpx<float>>(cpx<double> v) {
  return mk<float>((float)v.re, (float)v.im);
}
template <>
__device__ __forceinline__ cpx<double> narrow<cpx<double>>(cpx<double> v) {
  return v;
}
// load the running sum: exact ~= out - comp
template <typename V>
__device__ __forceinline__ wide_t<V> load_sum(const V *out, const V *comp,
                                              size_t o) {
  wide_t<V> s = widen(out[o]);
  if (comp) s = s - widen(comp[o]);
  return s;
}
// store the running sum (and its compensation if another launch follows)
template <typename V>
__device__ __forceinline__ void store_sum(V *out, V *comp, size_t o,
                                          wide_t<V> s, bool more) {
  const V r = narrow<V>(s);
  out[o] = r;
  if (more && comp) comp[o] = narrow<V>(widen(r) - s);
}

template <typename R>
__device__ __forceinline__ R fma_w(R w, R v, R acc) { return acc + w * v; }
template <typename R>
__device__ __forceinline__ cpx<R> fma_w(R w, cpx<R> v, cpx<R> acc) {
  return mk<R>(acc.re + w * v.re, acc.im + w * v.im);
}

// ---------------------------------------------------------------------------
// LDS-staged rotate-accumulate (_alg3d_bpp.py:73-81 rotates every y plane of
// the filtered slab by -angle and sums over angles).
//
// A block (4 waves) owns a 16x16 (z,x) output tile and YC consecutive y
// planes; wave w holds the 8x8 sub-tile (w>>1, w&1), its 16-lane group g the
// 4x4 quad (g>>1, g&1) of that, lane l of the group the voxel (l>>2, l&3).  Per angle the block copies a 32x32 window of each of its
// planes into LDS with direct global->LDS loads (16 bytes per lane when the
// row pitch allows it; no VGPR round trip) and every thread gathers its
// (ORDER+1)^2 taps from there.  k_rot_windows gives the window origin of
// each (angle, tile); the window holds every tap, mirrored ones included, of
// every in-bounds voxel of the tile.
// LDS layout: window row r takes 32*sizeof(V) bytes; its 16-byte granules are
// permuted g -> g ^ 2*(r mod 4).  The LDS serves 128 bytes per clock and
// resolves bank conflicts within groups of 16 lanes; a group holds a 4x4
// voxel quad, whose taps land on rotated 4x4 sample patterns, and the
// permutation spreads those over the 32 banks for all rotation angles
// (about 1.7 passes per 8-byte read on average instead of 4).
// Input coordinates are computed exactly as scipy does (float64, no FMA), so
// the out-of-bounds set and the tap positions are the reference's; the
// weights and tap sums of a float32 path are formed in float32 from the exact
// fractional offsets (each plane's sum widened exactly) and the sums over
// angles in float64; between launches the float32 total carries a
// compensation term.  Blocks are numbered so that all tiles of one y-chunk
// land on the same XCD (dispatch is round-robin over the 8 XCDs): they share
// the chunk's planes in that XCD's L2.
// ---------------------------------------------------------------------------
constexpr int kRotTile = 16;  // output tile edge (z and x)
constexpr int kRotWin = 32;   // LDS window edge per plane

typedef __attribute__((address_space(3))) void *lds_vptr;
typedef __attribute__((address_space(1))) void *gbl_vptr;

// s_waitcnt immediate (gfx9 encoding) waiting until at most n vector memory
// operations of this wave are outstanding; export and LDS counters untouched.
__host__ __device__ constexpr int waitcnt_vm(int n) {
  return (n & 15) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8);
}

// The same for a wave-uniform count known at run time (counts above 8 wait
// for at most 8 outstanding operations, which is stricter).
__device__ __forceinline__ void wait_vm(int n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(waitcnt_vm(0)); break;
    case 1: __builtin_amdgcn_s_waitcnt(waitcnt_vm(1)); break;
    case 2: __builtin_amdgcn_s_waitcnt(waitcnt_vm(2)); break;
    case 3: __builtin_amdgcn_s_waitcnt(waitcnt_vm(3)); break;
    case 4: __builtin_amdgcn_s_waitcnt(waitcnt_vm(4)); break;
    case 5: __builtin_amdgcn_s_waitcnt(waitcnt_vm(5)); break;
    case 6: __builtin_amdgcn_s_waitcnt(waitcnt_vm(6)); break;
    case 7: __builtin_amdgcn_s_waitcnt(waitcnt_vm(7)); break;
    default: __builtin_amdgcn_s_waitcnt(waitcnt_vm(8)); break;
  }
}

// First window row (or column) for the coordinate range [lo, hi] of a tile:
// the spline start of the lowest in-range coordinate, one sample lower for
// the zero-weight mirrored tap that odd orders touch at the upper edge,
// aligned down to `al` samples and clamped into the plane.  The taps of the
// tile then span at most 24 + ORDER samples (a 16-voxel tile edge covers at
// most 15*sqrt(2) input samples), which leaves room for the alignment.
template <int ORDER>
__device__ __forceinline__ int rot_origin(double lo, double hi, int n, int al) {
  lo = fmax(lo, 0.0);
  hi = fmin(hi, (double)(n - 1));
  int w = lo <= hi ? spline_start<ORDER>(lo) - 1 : 0;
  w = w < 0 ? 0 : w - w % al;
  const int wmax = n > kRotWin ? n - kRotWin : 0;
  return w < wmax ? w : wmax;
}

// Rows [lo, hi] (window coordinates, inclusive, clipped to the window) that
// hold every tap of an in-range coordinate of [clo, chi], mirrored taps
// included: from one below the lowest spline start to the highest start
// plus ORDER, within the plane (mirrored taps land inside that range).
// Empty (lo > hi) when no coordinate is in range.
template <int ORDER>
__device__ __forceinline__ int rot_extent(double clo, double chi, int n, int w0) {
  clo = fmax(clo, 0.0);
  chi = fmin(chi, (double)(n - 1));
  if (!(clo <= chi)) return 1;  // lo = 1, hi = 0
  int lo = spline_start<ORDER>(clo) - 1, hi = spline_start<ORDER>(chi) + ORDER;
  lo = (lo < 0 ? 0 : lo) - w0;
  hi = (hi > n - 1 ? n - 1 : hi) - w0;
  lo = lo < 0 ? 0 : lo;
  hi = hi > kRotWin - 1 ? kRotWin - 1 : hi;
  return lo | (hi << 8);
}

// Window origins (z, x) of every (angle, tile) and the extents of the
// window the tile's taps use (packed rows lo | hi << 8 | cols lo << 16 |
// hi << 24): the copy into LDS fetches only those.  A tile's computed input
// coordinates are monotone in each output index (every rounding step is), so
// the four corners, evaluated with the voxels' own float64 operations, span
// them exactly.
template <int ORDER>
__global__ void k_rot_windows(const double *__restrict__ params, int nb, int ntz,
                              int ntx, int nz, int nx, int al,
                              int4 *__restrict__ win) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nt = ntz * ntx;
  if (i >= (long long)nb * nt) return;
  const int a = (int)(i / nt), ti = (int)(i - (long long)a * nt);
  const int tz = ti / ntx, tx = ti - tz * ntx;
  const double *p = params + 6 * (size_t)a;
  const double z0 = tz * kRotTile, z1 = min(tz * kRotTile + kRotTile - 1, nz - 1);
  const double x0 = tx * kRotTile, x1 = min(tx * kRotTile + kRotTile - 1, nx - 1);
  double zlo = 0, zhi = 0, xlo = 0, xhi = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double oz = (c & 1) ? z1 : z0, ox = (c & 2) ? x1 : x0;
    const double cz = affine2(oz, ox, p[0], p[1], p[4]);
    const double cx = affine2(oz, ox, p[2], p[3], p[5]);
    zlo = c ? fmin(zlo, cz) : cz;
    zhi = c ? fmax(zhi, cz) : cz;
    xlo = c ? fmin(xlo, cx) : cx;
    xhi = c ? fmax(xhi, cx) : cx;
  }
  const int wz0 = rot_origin<ORDER>(zlo, zhi, nz, 1);
  const int wx0 = rot_origin<ORDER>(xlo, xhi, nx, al);
  const int ez = rot_extent<ORDER>(zlo, zhi, nz, wz0);
  const int ex = rot_extent<ORDER>(xlo, xhi, nx, wx0);
  win[i] = make_int4(wz0, wx0, ez | (ex << 16), 0);
}

template <typename R>
__device__ __forceinline__ R mul_w(R w, R v) { return w * v; }
template <typename R>
__device__ __forceinline__ cpx<R> mul_w(R w, cpx<R> v) {
  return mk<R>(w * v.re, w * v.im);
}

template <typename R, typename V, int ORDER, int YC, int VEC>
struct RotTile {
  static constexpr int ELB = (int)sizeof(V);     // bytes per sample
  static constexpr int ROWB = kRotWin * ELB;     // bytes per window row
  static constexpr int NG = ROWB / 16;           // 16-byte granules per row
  static constexpr int AL = 16 / ELB;            // samples per granule
  static constexpr int PLB = kRotWin * ROWB;     // bytes per plane window
  static constexpr int TB = YC * PLB;            // bytes per tile buffer
  static constexpr int NW = 4;                   // waves per block
  static constexpr int BPI = 256 * VEC;          // bytes per wave load
  static constexpr int K = TB / BPI / NW;        // loads per wave per angle
  static constexpr int RPI = BPI / ROWB;         // rows per wave load (VEC 4)
  static constexpr int PL = RPI > 0 ? kRotWin / RPI : 1;  // loads per plane (VEC 4)
  // 16-byte loads whose RPI rows start at multiples of 4 (samples of at
  // most 8 bytes): window rows are placed densely from the extent on
  static constexpr bool DENSE = VEC == 4 && RPI >= 4 && RPI % 4 == 0;
  static constexpr int NLO = DENSE ? 2 : (VEC == 4 ? (NG / RPI < K ? NG / RPI : K) : 1);
  static_assert(ELB <= 16 && 16 % ELB == 0 && NG >= 8, "sample size");
  static_assert(TB % (BPI * NW) == 0, "tile must split evenly");
  static_assert(K <= 63, "vmcnt range");
  static_assert(VEC == 1 || (RPI >= 1 && NG % RPI == 0), "row split");

  // granule permutation of window row r
  __device__ static __forceinline__ unsigned swz(int r) {
    return ((unsigned)r & 3u) << 1;
  }

  // Per-lane constants of the 16-byte loads: lane l covers row l / NG of a
  // load's RPI rows and granule position l mod NG, which holds the plane's
  // granule (l mod NG) ^ swz(row).  DENSE: loads start at rows that are
  // multiples of 4, so that granule depends on the lane only; lo[0] is the
  // row's byte offset in the plane, lo[1] the granule.  Otherwise lo[i] is
  // the lane's byte offset in load i (they repeat every NLO loads).
  __device__ static __forceinline__ void lane_offsets(unsigned *lo, int wave,
                                                      int lane, int nx) {
    if (VEC != 4) return;
    const int rl = lane / NG, g = lane % NG;
    if constexpr (DENSE) {
      lo[0] = (unsigned)(rl * nx) * ELB;
      lo[1] = (unsigned)g ^ swz(rl);
    } else {
#pragma unroll
      for (int i = 0; i < NLO; ++i) {
        const int r = (wave * K + i) * RPI + rl;
        lo[i] = (unsigned)(rl * nx + (int)(g ^ swz(r)) * AL) * ELB;
      }
    }
  }

  // Copy the windows of the block's planes for one angle into dst with
  // direct global->LDS loads, 4*VEC bytes per lane, fetching only the window
  // extent `ext` (see k_rot_windows).  16-byte loads are used only when both
  // plane edges are >= 32, so windows lie inside the plane.  DENSE: the
  // block's loads of a plane cover RPI consecutive rows each, from the
  // extent's first row (aligned down to 4) on, the loads past the extent's
  // last row repeating its last row block (same source and destination, no
  // new cache lines); a lane whose granule lies outside the extent's
  // columns fetches the nearest one inside instead (a slot no tap reads; the
  // same cache lines).  The global address is a uniform base plus a fixed
  // lane offset.  Every window row from the extent's first on then holds
  // plane samples; rows above it are never read (the stand-in taps of
  // voxels outside the plane use the first extent row, see prep).
  // Otherwise load k (= wave*K + i) of the block fills tile bytes
  // [k*BPI, (k+1)*BPI), and slots outside the extent load the block's first
  // sample(s) instead (one cache line for the whole wave), so every slot
  // holds a finite value.  4-byte loads also clamp rows and columns into
  // the plane (slots outside it receive samples that no tap reads).  Every
  // load instruction is issued, so the vmcnt waits keep counting K per wave.
  // The two forms are separate instantiations: a branch between them lets
  // the compiler merge their last loads, which loses the LDS object of the
  // merged load and makes every later LDS read wait for all loads in flight.
  __device__ static __forceinline__ int issue(const V *coefA, int wz0, int wx0,
                                              int ext, char *dst, size_t plane,
                                              int nyv, int nz, int nx, int wave,
                                              int lane, const unsigned *lo) {
    const char *src = reinterpret_cast<const char *>(coefA);
    // the mask bounds the destination offsets for the compiler: it then
    // knows these stores stay inside dst and does not wait for them before
    // reads of another buffer
    wave &= NW - 1;
    const int rlo = ext & 0xff, rhi = (ext >> 8) & 0xff;
    const int clo = (ext >> 16) & 0xff, chi = (ext >> 24) & 0xff;
    if constexpr (DENSE) {
      const Dense d = dense(wave, nyv, plane);
      return issue_dense(src, wz0, wx0, ext, dst, d, nx, lo);
    } else if constexpr (VEC == 4) {
      const unsigned uo = (unsigned)(wz0 * nx + wx0) * ELB;
      const int rl = lane / NG, gl = lane % NG;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int k = wave * K + i;
        const int rt = k * RPI;
        const int q = rt / kRotWin, r = rt % kRotWin;
        const int qq = q < nyv ? q : 0;
        const char *g = src + ((size_t)qq * plane + (size_t)r * nx) * ELB;
        const int row = r + rl;
        const int c0 = (int)(gl ^ swz(rt + rl)) * AL;
        const bool need = row >= rlo && row <= rhi && c0 <= chi && c0 + AL - 1 >= clo;
        const char *a = need ? g + (lo[i % NLO] + uo) : src;
        __builtin_amdgcn_global_load_lds((gbl_vptr)a, (lds_vptr)(dst + k * BPI), 16,
                                         0, 0);
      }
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int k = wave * K + i;
        const int b = k * BPI + lane * 4 * VEC;
        const int rt = b / ROWB, wb = b % ROWB;
        const int q = rt / kRotWin, r = rt % kRotWin;
        const int sb = (int)(((unsigned)(wb >> 4) ^ swz(r)) << 4) + (wb & 15);
        const int e = sb / ELB, eb = sb % ELB;
        int row = wz0 + r;
        row = row < nz ? row : nz - 1;
        int col = wx0 + e;
        col = col < nx - 1 ? col : nx - 1;
        const int qq = q < nyv ? q : 0;
        const bool need = r >= rlo && r <= rhi && e >= clo && e <= chi;
        // the stand-in keeps the byte position within the sample, so a
        // skipped slot holds the bytes of one whole (finite) sample
        const char *g =
            need ? src + ((size_t)qq * plane + (size_t)row * nx + col) * ELB + eb
                 : src + eb;
        __builtin_amdgcn_global_load_lds((gbl_vptr)g, (lds_vptr)(dst + k * BPI),
                                         4, 0, 0);
      }
    }
    return K;
  }

  // DENSE copies: every load of a wave fills rows of one plane q; its byte
  // offset from the angle's first plane (plane 0 when q does not exist),
  // its LDS offset and the wave's first load index within the plane.
  // Loop-invariant: kernels form it once per buffer.
  struct Dense {
    size_t qoff;
    int dq, kkb;
  };
  static_assert(!DENSE || PL % K == 0, "a wave's loads stay in one plane");
  __device__ static __forceinline__ Dense dense(int wave, int nyv, size_t plane) {
    wave &= NW - 1;
    const int q = wave * K / PL;
    Dense d;
    d.qoff = (size_t)(q < nyv ? q : 0) * plane * ELB;
    d.dq = q * PLB;
    d.kkb = wave * K % PL;
    return d;
  }

  // Returns the number of loads this wave issued (wave-uniform, 0..K).
  __device__ static __forceinline__ int issue_dense(const char *src, int wz0, int wx0,
                                                    int ext, char *dst, const Dense &d,
                                                    int nx, const unsigned *lo) {
    const int rlo = ext & 0xff, rhi = (ext >> 8) & 0xff;
    const int clo = (ext >> 16) & 0xff, chi = (ext >> 24) & 0xff;
    const unsigned g0 = (unsigned)clo / AL, g1 = (unsigned)chi / AL;
    unsigned gc = lo[1] < g0 ? g0 : lo[1];
    gc = gc > g1 ? g1 : gc;
    const unsigned voff = lo[0] + gc * 16u;
    const int rb0 = (rlo & ~3) + RPI * d.kkb;
    // this wave's loads that start at or before the extent's last row (and
    // inside the window)
    int last = rhi < kRotWin - 1 ? rhi : kRotWin - 1;
    last = last >= rlo ? last : (rlo & ~3);  // empty extent: one row block
    int n = (last - rb0) / RPI + 1;
    n = rb0 > last ? 0 : (n < K ? n : K);
    const char *wbase = src + d.qoff + (size_t)(unsigned)(wz0 * nx + wx0) * ELB;
    char *dq = dst + d.dq;
    const unsigned rowb = (unsigned)nx * ELB;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i < n) {
        // a block past the window's last RPI rows starts there instead
        // (RPI = 8: overlapping rows get the same samples twice)
        int rb = rb0 + RPI * i;
        rb = rb < kRotWin - RPI ? rb : kRotWin - RPI;
        __builtin_amdgcn_global_load_lds(
            (gbl_vptr)(wbase + (size_t)((unsigned)rb * rowb) + voff),
            (lds_vptr)(dq + rb * ROWB), 16, 0, 0);
      }
    }
    return n;
  }

  __device__ static __forceinline__ V tap(const char *tile, unsigned off) {
    return *reinterpret_cast<const V *>(tile + off);
  }

  // Window row of tap i for spline start s (plane edge n, window origin w0),
  // mirrored as scipy's mirror mode does.  With 16-byte loads both plane
  // edges are >= 32: taps of an in-range coordinate then leave the plane by
  // less than n - 1 samples, where one reflection is the whole mirror rule,
  // and below the centre tap only the lower edge can be crossed, above it
  // only the upper one.  The 4-byte variant (any plane size) takes the
  // general rule.
  template <int I>
  __device__ static __forceinline__ int tap_row(int s, int n, int w0) {
    constexpr int C = ORDER / 2;  // centre tap of the spline support
    const int t = s + I;
    if constexpr (VEC == 1) return mirror_idx(t, n) - w0;
    if (I < C) return (t < 0 ? -t : t) - w0;
    if (I > C) {
      const int d = n - 1 - t;
      return (n - 1 - w0) - (d < 0 ? -d : d);
    }
    return t - w0;
  }

  // Tap offsets (window bytes) and weights of one voxel for one angle.
  struct Taps {
    unsigned off[ORDER + 1][ORDER + 1];
    R wf[ORDER + 1][ORDER + 1];
  };

  // Gather one angle for the voxel (oz, ox) of every plane and add it to acc.
  // Branch-free: a voxel outside the rotated plane (scipy's constant mode,
  // value 0) or outside the volume gets zero weights and reads in-window
  // samples.
  __device__ static __forceinline__ void compute(const char *tile, int wz0,
                                                 int wx0, int ext, const double *p,
                                                 double oz, double ox, int nz,
                                                 int nx, bool own,
                                                 wide_t<V> *acc) {
    Taps t;
    prep(wz0, wx0, ext, p, oz, ox, nz, nx, own, t);
    gather(tile, t, acc);
  }

  // The tap set of the voxel (oz, ox) for one angle (coordinates, validity,
  // weights and window offsets; the same for every plane of the window).
  // A voxel outside the rotated plane reads the taps from the window's first
  // extent row on (zero weights; rows from there on hold plane samples).
  __device__ static __forceinline__ void prep(int wz0, int wx0, int ext,
                                              const double *p, double oz, double ox,
                                              int nz, int nx, bool own, Taps &t) {
    const double cz = affine2(oz, ox, p[0], p[1], p[4]);
    const double cx = affine2(oz, ox, p[2], p[3], p[5]);
    const bool valid = own && cz >= 0.0 && cz <= (double)(nz - 1) && cx >= 0.0 &&
                       cx <= (double)(nx - 1);
    const double fz = (ORDER & 1) ? floor(cz) : floor(cz + 0.5);
    const double fx = (ORDER & 1) ? floor(cx) : floor(cx + 0.5);
    const int sz = valid ? (int)fz - ORDER / 2 : wz0 + (ext & 0xff);
    const int sx = valid ? (int)fx - ORDER / 2 : wx0 + 1;
    // coordinates and spline start in float64; the weights of the float32
    // path in float32 from the exact fractional offsets, their products
    // formed once for both planes
    double wzd[ORDER + 1], wxd[ORDER + 1];
    R wzr[ORDER + 1], wxr[ORDER + 1];
    if constexpr (sizeof(R) == 4) {
      // the fractional offsets are exact in float64; the weights in float32
      spline_weights_t<ORDER, R>((R)(cz - fz), wzr);
      spline_weights_t<ORDER, R>((R)(cx - fx), wxr);
    } else {
      spline_weights_t<ORDER, double>(cz - fz, wzd);
      spline_weights_t<ORDER, double>(cx - fx, wxd);
#pragma unroll
      for (int i = 0; i <= ORDER; ++i) {
        wzr[i] = (R)wzd[i];
        wxr[i] = (R)wxd[i];
      }
    }
#pragma unroll
    for (int i = 0; i <= ORDER; ++i) wzr[i] = valid ? wzr[i] : R(0);
    int zr[ORDER + 1], xc[ORDER + 1];
    rows_of<0>(sz, nz, wz0, zr);
    rows_of<0>(sx, nx, wx0, xc);
    unsigned rb[ORDER + 1], gs[ORDER + 1], cb[ORDER + 1];
#pragma unroll
    for (int i = 0; i <= ORDER; ++i) {
      rb[i] = (unsigned)zr[i] * ROWB;
      gs[i] = swz(zr[i]) << 4;
      cb[i] = (unsigned)xc[i] * ELB;
    }
#pragma unroll
    for (int i = 0; i <= ORDER; ++i)
#pragma unroll
      for (int j = 0; j <= ORDER; ++j) {
        t.wf[i][j] = wzr[i] * wxr[j];
        t.off[i][j] = rb[i] + (cb[j] ^ gs[i]);
      }
  }

  // Tap sums in the path's precision, each plane's sum widened into the
  // float64 angle sum.
  __device__ static __forceinline__ void gather(const char *tile, const Taps &t,
                                                wide_t<V> *acc) {
#pragma unroll
    for (int q = 0; q < YC; ++q) {
      V v = mul_w(t.wf[0][0], tap(tile, t.off[0][0] + q * PLB));
#pragma unroll
      for (int i = 0; i <= ORDER; ++i)
#pragma unroll
        for (int j = 0; j <= ORDER; ++j)
          if (i | j) v = fma_w(t.wf[i][j], tap(tile, t.off[i][j] + q * PLB), v);
      acc[q] = acc[q] + widen(v);
    }
  }

  template <int I>
  __device__ static __forceinline__ void rows_of(int s, int n, int w0, int *r) {
    r[I] = tap_row<I>(s, n, w0);
    if constexpr (I < ORDER) rows_of<I + 1>(s, n, w0, r);
  }
};

// s_waitcnt immediate (gfx9 encoding) waiting until no LDS (and scalar
// memory) operation of this wave is outstanding; vector memory and export
// counters untouched.
__host__ __device__ constexpr int waitcnt_lgkm0() {
  return 0xF | (0x7 << 4) | (0x0 << 8) | (0x3 << 14);
}

// Workgroup barrier without the memory fence of __syncthreads(): that fence
// would also wait for the prefetch of the next tile.  Ordering is explicit
// instead: each wave waits (vmcnt) for its own loads of the tile it is about
// to read, and, before the barrier, for its LDS reads (lgkmcnt): a read
// still in flight at the barrier could be serviced after another wave's
// direct global->LDS load has refilled the buffer behind it (issuing the
// reads before the barrier is not enough; the compiler consumes their
// results later).  The signal fences stop the compiler from moving memory
// operations across the barrier.
__device__ __forceinline__ void barrier_no_fence() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(waitcnt_lgkm0());
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Zero a block's LDS window buffers once before the first window copy:
// the copies fetch only each angle's window extent, so a slot outside it
// then holds zero or a sample of an earlier angle, both finite, for the
// zero-weight taps of voxels outside the rotated plane.
__device__ __forceinline__ void zero_lds(char *buf, int bytes) {
  for (int o = (int)threadIdx.x * 16; o < bytes; o += (int)blockDim.x * 16)
    *reinterpret_cast<int4 *>(buf + o) = make_int4(0, 0, 0, 0);
  __builtin_amdgcn_s_waitcnt(waitcnt_lgkm0());
  __builtin_amdgcn_s_barrier();
}

// Order in which the blocks of one y-chunk take the tiles: along a Z-order
// curve for square power-of-two tile grids, else in 4 x 4 groups of tiles
// (row-major groups) where the grid allows, else row-major.  Tiles whose
// windows overlap then run close together in dispatch order, so a window
// line one block fetched is still in the XCD's L2 when its vertical
// neighbours ask for it (row-major order puts them a whole tile row apart).
// Measured at 256^2 x 360, rotate stage per reconstruction: row-major 15.3
// ms, 4 x 4 groups 14.6, Z-order 14.4 (48-angle launches).
__device__ __forceinline__ int tile_order(int t, int ntx, int ntz) {
  if (ntx == ntz && (ntx & (ntx - 1)) == 0) {
    int tx = 0, tz = 0;
    for (int b = 0; (1 << b) < ntx; ++b) {
      tx |= ((t >> (2 * b)) & 1) << b;
      tz |= ((t >> (2 * b + 1)) & 1) << b;
    }
    return tz * ntx + tx;
  }
  if (ntx % 4 == 0 && ntz % 4 == 0) {
    const int g = t >> 4, w = t & 15;
    const int gx = g % (ntx / 4), gz = g / (ntx / 4);
    return (gz * 4 + (w >> 2)) * ntx + gx * 4 + (w & 3);
  }
  return t;
}

// ag > 0: angle groups along blockIdx.y; group gy takes angles
// [gy*ag, gy*ag + ag) and writes its own running sum and compensation at
// out/comp + gy*gstride (k_fold_groups adds the groups up).
template <typename R, typename V, int ORDER, int YC, int VEC>
__global__ void __launch_bounds__(256)
    k_rotate_lds(const V *__restrict__ coef, int nb,
                 const double *__restrict__ params,
                 const int4 *__restrict__ win, V *__restrict__ out,
                 V *__restrict__ comp, int ny, int nz, int nx, int first,
                 int last, int ntx, int ntz, int xcd_map, int ag,
                 size_t gstride) {
  using RT = RotTile<R, V, ORDER, YC, VEC>;
  if (ag > 0) {
    const int a0 = blockIdx.y * ag;
    nb = min(ag, nb - a0);
    coef += (size_t)a0 * ny * nz * nx;
    params += 6 * (size_t)a0;
    win += (size_t)a0 * ntx * ntz;
    out += blockIdx.y * gstride;
    comp += blockIdx.y * gstride;
    first = 1;
    last = 0;
  }
  // two tile buffers as separate LDS objects: reads of one are not ordered
  // behind the in-flight direct loads into the other
  __shared__ __align__(16) char bufA[RT::TB];
  __shared__ __align__(16) char bufB[RT::TB];
  const int nt = ntx * ntz;
  const int nyc = (ny + YC - 1) / YC;
  int yc, ti;
  if (xcd_map) {
    const int bid = blockIdx.x;
    yc = (bid & 7) + 8 * (bid / (8 * nt));
    ti = tile_order((bid >> 3) % nt, ntx, ntz);
    if (yc >= nyc) return;  // uniform per block
  } else {
    yc = blockIdx.x / nt;
    ti = blockIdx.x - yc * nt;
  }
  const int tz = ti / ntx, tx = ti - tz * ntx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int z = tz * kRotTile + (wave >> 1) * 8 + ((lane >> 5) << 2) + ((lane >> 2) & 3);
  const int x = tx * kRotTile + (wave & 1) * 8 + (((lane >> 4) & 1) << 2) + (lane & 3);
  const int y0 = yc * YC;
  const int nyv = ny - y0 < YC ? ny - y0 : YC;
  const bool own = x < nx && z < nz;
  const size_t plane = (size_t)nz * nx;
  unsigned lo[RT::NLO];
  RT::lane_offsets(lo, wave, lane, nx);
  wide_t<V> sum[YC];
#pragma unroll
  for (int q = 0; q < YC; ++q) {
    acc_zero(sum[q]);
    if (!first && own && q < nyv)
      sum[q] = load_sum(out, comp, ((size_t)z * ny + y0 + q) * nx + x);
  }
  const double oz = (double)z, ox = (double)x;
  const V *base = coef + (size_t)y0 * plane;
  const size_t aStride = (size_t)ny * plane;
  const int4 *wt = win + ti;
  zero_lds(bufA, RT::TB);
  zero_lds(bufB, RT::TB);
  int4 wA = make_int4(0, 0, 1, 0), wB = make_int4(0, 0, 1, 0);
  // Software pipeline over angles, two per iteration, every prefetch issued
  // from the loop body: angle a+1 goes to B while a (in A) is computed, then
  // a+2 goes to A while a+1 (in B) is computed.  The first iteration starts
  // at a = -1 with nothing to compute.
  for (int a = -1; a < nb; a += 2) {
    if (a + 1 < nb) {
      wB = wt[(size_t)(a + 1) * nt];
      wait_vm(RT::issue(base + (a + 1) * aStride, wB.x, wB.y, wB.z, bufB, plane, nyv,
                        nz, nx, wave, lane, lo));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    barrier_no_fence();
    if (a >= 0)
      RT::compute(bufA, wA.x, wA.y, wA.z, params + 6 * a, oz, ox, nz, nx, own, sum);
    barrier_no_fence();
    if (a + 1 < nb) {
      if (a + 2 < nb) {
        wA = wt[(size_t)(a + 2) * nt];
        wait_vm(RT::issue(base + (a + 2) * aStride, wA.x, wA.y, wA.z, bufA, plane, nyv,
                          nz, nx, wave, lane, lo));
      } else {
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
      }
      barrier_no_fence();
      RT::compute(bufB, wB.x, wB.y, wB.z, params + 6 * (a + 1), oz, ox, nz, nx, own,
                  sum);
      barrier_no_fence();
    }
  }
  if (own) {
#pragma unroll
    for (int q = 0; q < YC; ++q) {
      if (q >= nyv) break;
      store_sum(out, comp, ((size_t)z * ny + y0 + q) * nx + x, sum[q], !last);
    }
  }
}

// Adds the G angle-group sums of k_rotate_lds (sum g at part + g*n, its
// compensation at part + (G + g)*n) to the running sum in out/comp, in group
// order, in float64.
template <typename V>
__global__ void __launch_bounds__(256)
    k_fold_groups(V *__restrict__ out, V *__restrict__ comp,
                  const V *__restrict__ part, int G, size_t n, int first,
                  int last) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wide_t<V> s;
  acc_zero(s);
  if (!first) s = load_sum(out, comp, i);
  for (int g = 0; g < G; ++g)
    s = s + (widen(part[(size_t)g * n + i]) - widen(part[(size_t)(G + g) * n + i]));
  store_sum(out, comp, i, s, !last);
}

// Four planes in two half buffers of two planes, pipelined at half-angle
// granularity: planes 2-3 of angle a load while planes 0-1 of angle a are
// gathered, planes 0-1 of angle a+1 while planes 2-3 of angle a are.  The
// voxel's tap set (coordinates, weights, offsets) is formed once per angle
// for all four planes.
template <typename R, typename V, int ORDER, int VEC>
__global__ void __launch_bounds__(256)
    k_rotate_lds2(const V *__restrict__ coef, int nb,
                  const double *__restrict__ params,
                  const int4 *__restrict__ win, V *__restrict__ out,
                  V *__restrict__ comp, int ny, int nz, int nx, int first,
                  int last, int ntx, int ntz, int xcd_map) {
  constexpr int YC = 4;
  using RT = RotTile<R, V, ORDER, 2, VEC>;
  // the halves as separate LDS objects: reads of one are not ordered behind
  // the in-flight direct loads into the other
  __shared__ __align__(16) char h0[RT::TB];
  __shared__ __align__(16) char h1[RT::TB];
  const int nt = ntx * ntz;
  const int nyc = (ny + YC - 1) / YC;
  int yc, ti;
  if (xcd_map) {
    const int bid = blockIdx.x;
    yc = (bid & 7) + 8 * (bid / (8 * nt));
    ti = tile_order((bid >> 3) % nt, ntx, ntz);
    if (yc >= nyc) return;  // uniform per block
  } else {
    yc = blockIdx.x / nt;
    ti = blockIdx.x - yc * nt;
  }
  const int tz = ti / ntx, tx = ti - tz * ntx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int z = tz * kRotTile + (wave >> 1) * 8 + ((lane >> 5) << 2) + ((lane >> 2) & 3);
  const int x = tx * kRotTile + (wave & 1) * 8 + (((lane >> 4) & 1) << 2) + (lane & 3);
  const int y0 = yc * YC;
  const int nyv = ny - y0 < YC ? ny - y0 : YC;
  // planes of each half that exist; an empty upper half reads the lower
  // half's first plane (in bounds, never stored)
  const int nyv0 = nyv < 2 ? nyv : 2, nyv1 = nyv > 2 ? nyv - 2 : 0;
  const bool own = x < nx && z < nz;
  const size_t plane = (size_t)nz * nx;
  unsigned lo[RT::NLO];
  RT::lane_offsets(lo, wave, lane, nx);
  wide_t<V> sum[YC];
#pragma unroll
  for (int q = 0; q < YC; ++q) {
    acc_zero(sum[q]);
    if (!first && own && q < nyv)
      sum[q] = load_sum(out, comp, ((size_t)z * ny + y0 + q) * nx + x);
  }
  const double oz = (double)z, ox = (double)x;
  const V *base0 = coef + (size_t)y0 * plane;
  const V *base1 = nyv1 > 0 ? base0 + 2 * plane : base0;
  const size_t aStride = (size_t)ny * plane;
  const int4 *wt = win + ti;
  typename RT::Dense d0{}, d1{};
  if constexpr (RT::DENSE) {
    d0 = RT::dense(wave, nyv0, plane);
    d1 = RT::dense(wave, nyv1, plane);
  }
  // window copy of one half (lower: planes 0-1, upper: 2-3)
  // window copy of one half (lower: planes 0-1, upper: 2-3); returns the
  // wave's load count
  auto copy = [&](const V *b, const int4 &w, char *dst, bool upper) {
    if constexpr (RT::DENSE)
      return RT::issue_dense(reinterpret_cast<const char *>(b), w.x, w.y, w.z, dst,
                             upper ? d1 : d0, nx, lo);
    else
      return RT::issue(b, w.x, w.y, w.z, dst, plane, upper ? nyv1 : nyv0, nz, nx, wave,
                       lane, lo);
  };
  zero_lds(h0, RT::TB);
  zero_lds(h1, RT::TB);
  int4 w = wt[0];  // nb >= 1 (the launcher checks)
  copy(base0, w, h0, false);
  // angles 0 .. nb-2 prefetch the next angle's lower half; the last angle
  // (peeled) drains instead
  for (int a = 0; a + 1 < nb; ++a) {
    // the loads just issued may stay in flight; the earlier ones have landed
    wait_vm(copy(base1 + a * aStride, w, h1, true));
    barrier_no_fence();
    typename RT::Taps t;
    RT::prep(w.x, w.y, w.z, params + 6 * a, oz, ox, nz, nx, own, t);
    RT::gather(h0, t, sum);
    barrier_no_fence();
    w = wt[(size_t)(a + 1) * nt];
    wait_vm(copy(base0 + (a + 1) * aStride, w, h0, false));
    barrier_no_fence();
    RT::gather(h1, t, sum + 2);
    barrier_no_fence();
  }
  {
    const int a = nb - 1;
    copy(base1 + a * aStride, w, h1, true);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    barrier_no_fence();
    typename RT::Taps t;
    RT::prep(w.x, w.y, w.z, params + 6 * a, oz, ox, nz, nx, own, t);
    RT::gather(h0, t, sum);
    RT::gather(h1, t, sum + 2);
  }
  if (own) {
#pragma unroll
    for (int q = 0; q < YC; ++q) {
      if (q >= nyv) break;
      store_sum(out, comp, ((size_t)z * ny + y0 + q) * nx + x, sum[q], !last);
    }
  }
}

// Whole-sample mirror of an index at most n - 1 outside [0, n).
__device__ __forceinline__ int mirror1(int idx, int n) {
  return idx < 0 ? -idx : (idx >= n ? 2 * n - 2 - idx : idx);
}

// ---------------------------------------------------------------------------
// affine gather-accumulate for the tilted axis (_alg3d_bppt.py:904-944):
// output voxel (z,y,x) is sample o = (x, Ny-1-y, z) of the (x, -y, z) frame;
// input index c = Dinv @ o + offset in that frame; coefficient (i0,i1,i2)
// of the frame lives at slab[y = Ny-1-i1][z = i2][x = i0].
// ---------------------------------------------------------------------------
template <typename R, typename V, int ORDER>
__global__ void __launch_bounds__(256)
    k_affine_acc(const V *__restrict__ coef, int nb,
                 const double *__restrict__ params, V *__restrict__ out,
                 V *__restrict__ comp, int ny, int nz, int nx, int first,
                 int last) {
  // a wavefront takes an 8 x 8 (x, z) patch of one y plane: its taps then
  // fall on a compact rotated patch of the slab (a few cache lines per
  // gather) for every rotation angle, where a 64-voxel x row would spread
  // over up to 64 lines near a quarter turn
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * 8 + (lane & 7);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z = blockIdx.z * 8 + (lane >> 3);
  if (x >= nx || y >= ny || z >= nz) return;
  const size_t o = ((size_t)z * ny + y) * nx + x;
  wide_t<V> sum;
  if (!first)
    sum = load_sum(out, comp, o);
  else
    acc_zero(sum);
  const double o0 = (double)x, o1 = (double)(ny - 1 - y), o2 = (double)z;
  const size_t vol = (size_t)ny * nz * nx;
  const bool fast = nx >= 4 && ny >= 4 && nz >= 4;
  for (int a = 0; a < nb; ++a) {
    const double *p = params + 12 * a;
    const double c0 = affine3(o0, o1, o2, p[0], p[1], p[2], p[9]);
    const double c1 = affine3(o0, o1, o2, p[3], p[4], p[5], p[10]);
    const double c2 = affine3(o0, o1, o2, p[6], p[7], p[8], p[11]);
    if (c0 < 0.0 || c0 > (double)(nx - 1) || c1 < 0.0 ||
        c1 > (double)(ny - 1) || c2 < 0.0 || c2 > (double)(nz - 1))
      continue;
    R w0[ORDER + 1], w1[ORDER + 1], w2[ORDER + 1];
    int i0[ORDER + 1];
    size_t i1[ORDER + 1], i2[ORDER + 1];
    spline_weights<ORDER, R>(c0, w0);
    spline_weights<ORDER, R>(c1, w1);
    spline_weights<ORDER, R>(c2, w2);
    const int s0 = spline_start<ORDER>(c0), s1 = spline_start<ORDER>(c1),
              s2 = spline_start<ORDER>(c2);
    // taps of an in-range coordinate leave the volume by at most
    // (ORDER + 1) / 2 <= 3 samples: with every edge >= 4 one reflection is
    // the whole mirror rule (no integer division)
#pragma unroll
    for (int j = 0; j <= ORDER; ++j) {
      i0[j] = fast ? mirror1(s0 + j, nx) : mirror_idx(s0 + j, nx);
      i1[j] = (size_t)(ny - 1 - (fast ? mirror1(s1 + j, ny) : mirror_idx(s1 + j, ny))) *
              nz * nx;
      i2[j] = (size_t)(fast ? mirror1(s2 + j, nz) : mirror_idx(s2 + j, nz)) * nx;
    }
    const V *base = coef + (size_t)a * vol;
    V v;
    acc_zero(v);
#pragma unroll
    for (int q = 0; q <= ORDER; ++q) {
      V tq;
      acc_zero(tq);
#pragma unroll
      for (int r = 0; r <= ORDER; ++r) {
        const V *row = base + i1[q] + i2[r];
        V t;
        acc_zero(t);
#pragma unroll
        for (int pp = 0; pp <= ORDER; ++pp) t = fma_w(w0[pp], row[i0[pp]], t);
        tq = fma_w(w2[r], t, tq);
      }
      v = fma_w(w1[q], tq, v);
    }
    sum = sum + widen(v);
  }
  store_sum(out, comp, o, sum, !last);
}

// ---------------------------------------------------------------------------
// LDS-staged affine gather-accumulate for the tilted axis (same arithmetic
// as k_affine_acc).  A block (4 waves) owns an 8x8x8 output tile; wave w
// holds the (x, z) patch of the planes y0 + 2w and y0 + 2w + 1, in 4x4
// quads of 16 lanes.  Per angle the block copies
// the window of the slab that holds every tap (mirrored taps included) of
// the tile's in-range voxels into LDS -- at most WIN = 15 + ORDER samples
// per axis: a rotation moves a 7-sample tile edge over at most 7*sqrt(3)
// input samples along any axis -- and every thread gathers its taps from
// there.  k_affine_windows gives each (angle, tile) its window; blocks are
// numbered so that each XCD takes a contiguous run of tiles (neighbouring
// windows overlap and meet in that XCD's L2).
// ---------------------------------------------------------------------------
constexpr int kAffTile = 8;

template <int ORDER>
__host__ __device__ constexpr int aff_win() { return 15 + ORDER; }

// LDS pitches of the window box (elements): rows of 20 and planes a
// multiple of 16 spread a wave's taps -- a rotated 4x4 voxel patch per 16
// lanes -- over the banks (a bank model over the tilted rotations puts the
// 8-byte reads at 1.7 passes per 16 lanes, against 3.9 for the dense W x W
// planes).  16-byte elements keep dense planes: the padding would halve the
// blocks per CU.
template <int ORDER, typename V>
__host__ __device__ constexpr int aff_row_pitch() {
  return sizeof(V) >= 16 || aff_win<ORDER>() > 20 ? aff_win<ORDER>() : 20;
}
template <int ORDER, typename V>
__host__ __device__ constexpr int aff_plane_pitch() {
  return sizeof(V) >= 16 ? aff_win<ORDER>() * aff_win<ORDER>()
                         : (aff_win<ORDER>() * aff_row_pitch<ORDER, V>() + 15) / 16 * 16;
}

size_t affine_window_entries(int ny, int nz, int nx) {
  return (size_t)((nx + kAffTile - 1) / kAffTile) * ((ny + kAffTile - 1) / kAffTile) *
         ((nz + kAffTile - 1) / kAffTile);
}

// Window rows [lo, hi] along one axis (edge n) for coordinates spanning
// [clo, chi]: the taps start(c) .. start(c) + ORDER of the in-range part,
// after one whole-sample reflection at either end; packed lo | ext << 16 |
// direct << 24, ext 0 when no coordinate is in range, direct when no tap
// needs the reflection.
template <int ORDER>
__device__ __forceinline__ int aff_axis(double clo, double chi, int n) {
  clo = fmax(clo, 0.0);
  chi = fmin(chi, (double)(n - 1));
  if (!(clo <= chi)) return 0;
  const int s0 = spline_start<ORDER>(clo), s1 = spline_start<ORDER>(chi) + ORDER;
  int lo = s0 < 2 * n - 2 - s1 ? s0 : 2 * n - 2 - s1;
  int hi = s1 > -s0 ? s1 : -s0;
  lo = lo < 0 ? 0 : lo;
  hi = hi > n - 1 ? n - 1 : hi;
  int ext = hi - lo + 1;
  // every tap of the range inside [0, n) and inside the window: no mirror
  const int direct = s0 >= 0 && s1 <= n - 1 && ext <= aff_win<ORDER>() ? 1 : 0;
  ext = ext < aff_win<ORDER>() ? ext : aff_win<ORDER>();
  return lo | (ext << 16) | (direct << 24);
}

// Window of every (angle, tile) in the (x, -y, z) frame: the eight tile
// corners through the voxels' own float64 operations (each coordinate is
// monotone in each output index) bound the coordinates of the tile.
template <int ORDER>
__global__ void k_affine_windows(const double *__restrict__ params, int nb, int ntx,
                                 int nty, int ntz, int ny, int nz, int nx,
                                 int4 *__restrict__ win) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nt = ntx * nty * ntz;
  if (i >= (long long)nb * nt) return;
  const int a = (int)(i / nt), ti = (int)(i - (long long)a * nt);
  const int tx = ti % ntx, ty = (ti / ntx) % nty, tz = ti / (ntx * nty);
  const double *p = params + 12 * (size_t)a;
  const int x0 = tx * kAffTile, y0 = ty * kAffTile, z0 = tz * kAffTile;
  const int x1 = min(x0 + kAffTile - 1, nx - 1), y1 = min(y0 + kAffTile - 1, ny - 1),
            z1 = min(z0 + kAffTile - 1, nz - 1);
  double lo[3], hi[3];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double o0 = (c & 1) ? x1 : x0;
    const double o1 = (c & 2) ? (double)(ny - 1 - y0) : (double)(ny - 1 - y1);
    const double o2 = (c & 4) ? z1 : z0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = affine3(o0, o1, o2, p[3 * d], p[3 * d + 1], p[3 * d + 2], p[9 + d]);
      lo[d] = c ? fmin(lo[d], v) : v;
      hi[d] = c ? fmax(hi[d], v) : v;
    }
  }
  const int w0 = aff_axis<ORDER>(lo[0], hi[0], nx);
  const int w1 = aff_axis<ORDER>(lo[1], hi[1], ny);
  const int w2 = aff_axis<ORDER>(lo[2], hi[2], nz);
  const int inner = (w0 >> 24) & (w1 >> 24) & (w2 >> 24) & 1;
  win[i] = make_int4(w0 & 0xffff, w1 & 0xffff, w2 & 0xffff,
                     ((w0 >> 16) & 0xff) | (((w1 >> 16) & 0xff) << 8) |
                         (((w2 >> 16) & 0xff) << 16) | (inner << 24));
}

template <typename R, typename V, int ORDER>
__global__ void __launch_bounds__(256)
    k_affine_lds(const V *__restrict__ coef, int nb, const double *__restrict__ params,
                 const int4 *__restrict__ win, V *__restrict__ out,
                 V *__restrict__ comp, int ny, int nz, int nx, int first, int last,
                 int ntx, int nty, int ntz, int xcd_map) {
  constexpr int W = aff_win<ORDER>();
  constexpr int RP = aff_row_pitch<ORDER, V>(), PP = aff_plane_pitch<ORDER, V>();
  constexpr int NS = (W * W * W + 255) / 256;  // window samples per thread
  __shared__ V buf[W * PP + 1];  // + a sink slot for idle lanes
  const int nt = ntx * nty * ntz;
  int ti = blockIdx.x;
  if (xcd_map) {
    // XCD x (= block id mod 8) takes tiles [x * nt / 8, (x + 1) * nt / 8)
    const int per = nt / 8;
    ti = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int tx = ti % ntx, ty = (ti / ntx) % nty, tz = ti / (ntx * nty);
  // lanes 16g .. 16g + 15 take the 4x4 (x, z) quad g of the 8x8 patch
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = tx * kAffTile + 4 * ((lane >> 4) & 1) + (lane & 3),
            z = tz * kAffTile + 4 * (lane >> 5) + ((lane >> 2) & 3);
  const int ya = ty * kAffTile + 2 * wave;
  bool own[2];
  size_t o[2];
  wide_t<V> sum[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int y = ya + k;
    own[k] = x < nx && z < nz && y < ny;
    o[k] = ((size_t)z * ny + (own[k] ? y : 0)) * nx + x;
    acc_zero(sum[k]);
    if (!first && own[k]) sum[k] = load_sum(out, comp, o[k]);
  }
  const size_t vol = (size_t)ny * nz * nx;
  // The window as a W x W x W box in LDS, slot j * PP + k * RP + i =
  // sample (lo0 + i, lo1 + j, lo2 + k).  Per angle only the used extent
  // n0 x n1 x n2 is copied, enumerated densely (i fastest, so consecutive
  // threads read consecutive samples): element e has i = e mod n0,
  // k = (e / n0) mod n2, j = e / (n0 n2), the divisions as products with
  // float32 reciprocals from e + 1/2 (exact for the extents here: the
  // product stays far from the integers, rcp's error included).  Rounds of
  // 256 elements run while any element is left (a block-uniform test); in
  // the last round the idle lanes load the last element and store it to the
  // sink slot.  A thread issues all its loads before it stores any; slots
  // outside the extent keep older values that no tap reads.
  for (int a = 0; a < nb; ++a) {
    const int4 e = win[(size_t)a * nt + ti];
    const int lo0 = e.x, lo1 = e.y, lo2 = e.z;
    const int n0 = e.w & 0xff, n1 = (e.w >> 8) & 0xff, n2 = (e.w >> 16) & 0xff;
    const bool inner = (e.w >> 24) & 1;
    const V *org = coef + (size_t)a * vol + ((size_t)(ny - 1 - lo1) * nz + lo2) * nx + lo0;
    const int total = n0 * n1 * n2;
    if (total == 0) continue;  // no voxel of the tile in range (block-uniform)
    const float inv0 = __builtin_amdgcn_rcpf((float)n0), inv2 = __builtin_amdgcn_rcpf((float)n2);
    V vals[NS];
    int dst[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      if (256 * u < total) {
        const int e0 = (int)threadIdx.x + 256 * u;
        const bool in = e0 < total;
        const int el = in ? e0 : total - 1;
        const int r = (int)(((float)el + 0.5f) * inv0);
        const int i = el - __mul24(r, n0);
        const int j = (int)(((float)r + 0.5f) * inv2);
        const int k = r - __mul24(j, n2);
        vals[u] = org[(k - j * nz) * nx + i];
        dst[u] = in ? __mul24(j, PP) + __mul24(k, RP) + i : W * PP;
      }
    }
    __syncthreads();  // the previous angle's gathers are done
#pragma unroll
    for (int u = 0; u < NS; ++u)
      if (256 * u < total) buf[dst[u]] = vals[u];
    __syncthreads();
    const double *p = params + 12 * a;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double o0 = (double)x, o1 = (double)(ny - 1 - (ya + k)), o2 = (double)z;
      const double c0 = affine3(o0, o1, o2, p[0], p[1], p[2], p[9]);
      const double c1 = affine3(o0, o1, o2, p[3], p[4], p[5], p[10]);
      const double c2 = affine3(o0, o1, o2, p[6], p[7], p[8], p[11]);
      const bool valid = own[k] && !(c0 < 0.0 || c0 > (double)(nx - 1) || c1 < 0.0 ||
                                     c1 > (double)(ny - 1) || c2 < 0.0 ||
                                     c2 > (double)(nz - 1));
      if (!valid) continue;
      R w0[ORDER + 1], w1[ORDER + 1], w2[ORDER + 1];
      int i0[ORDER + 1], i1[ORDER + 1], i2[ORDER + 1];
      const int s0 = spline_taps<ORDER, R>(c0, w0), s1 = spline_taps<ORDER, R>(c1, w1),
                s2 = spline_taps<ORDER, R>(c2, w2);
      V v;
      acc_zero(v);
      if (inner) {
        // no tap leaves the volume: the taps sit at constant offsets from
        // the first one
        const V *b0 = buf + (s1 - lo1) * PP + (s2 - lo2) * RP + (s0 - lo0);
#pragma unroll
        for (int q = 0; q <= ORDER; ++q) {
          V tq;
          acc_zero(tq);
#pragma unroll
          for (int r = 0; r <= ORDER; ++r) {
            const V *row = b0 + q * PP + r * RP;
            V t;
            acc_zero(t);
#pragma unroll
            for (int pp = 0; pp <= ORDER; ++pp) t = fma_w(w0[pp], row[pp], t);
            tq = fma_w(w2[r], t, tq);
          }
          v = fma_w(w1[q], tq, v);
        }
      } else {
#pragma unroll
        for (int j = 0; j <= ORDER; ++j) {
          i0[j] = min(max(mirror1(s0 + j, nx) - lo0, 0), W - 1);
          i1[j] = min(max(mirror1(s1 + j, ny) - lo1, 0), W - 1) * PP;
          i2[j] = min(max(mirror1(s2 + j, nz) - lo2, 0), W - 1) * RP;
        }
#pragma unroll
        for (int q = 0; q <= ORDER; ++q) {
          V tq;
          acc_zero(tq);
#pragma unroll
          for (int r = 0; r <= ORDER; ++r) {
            const V *row = buf + i1[q] + i2[r];
            V t;
            acc_zero(t);
#pragma unroll
            for (int pp = 0; pp <= ORDER; ++pp) t = fma_w(w0[pp], row[i0[pp]], t);
            tq = fma_w(w2[r], t, tq);
          }
          v = fma_w(w1[q], tq, v);
        }
      }
      sum[k] = sum[k] + widen(v);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (own[k]) store_sum(out, comp, o[k], sum[k], !last);
}

// ---------------------------------------------------------------------------
// In-place rotation with scipy's aliasing for spline orders 0 and 1.
// The reference rotates each y-plane with
// scipy.ndimage.rotate(arr[...], output=arr[...]) (_alg3d_bpp.py:73-81).
// For order < 2 scipy does not prefilter, so it reads the array it is
// writing: output points are produced in C order (z outer, x inner) and later
// points see the already rotated values of earlier ones.  Every
// (angle, y-plane, real/imag part) is one independent sequential job here,
// replaying that order exactly on its own slab plane.
// ---------------------------------------------------------------------------
template <typename R, int ORDER>
__global__ void k_seq_rotate(R *data, long long njobs, int ny, int nz, int nx,
                             int ec, const double *__restrict__ params) {
  const long long job = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (job >= njobs) return;
  const long long a = job / ((long long)ny * ec);
  const int rem = (int)(job - a * ny * ec);
  const int y = rem / ec, comp = rem - y * ec;
  R *pl = data + ((a * ny + y) * (long long)nz * nx) * ec + comp;
  const double *p = params + 6 * a;
  for (int z = 0; z < nz; ++z) {
    for (int x = 0; x < nx; ++x) {
      const double cz = affine2((double)z, (double)x, p[0], p[1], p[4]);
      const double cx = affine2((double)z, (double)x, p[2], p[3], p[5]);
      R v = R(0);
      if (!(cz < 0.0 || cz > (double)(nz - 1) || cx < 0.0 ||
            cx > (double)(nx - 1))) {
        double wz[ORDER + 1], wx[ORDER + 1];
        spline_weights<ORDER, double>(cz, wz);
        spline_weights<ORDER, double>(cx, wx);
        const int sz = spline_start<ORDER>(cz), sx = spline_start<ORDER>(cx);
        double acc = 0.0;
        for (int i = 0; i <= ORDER; ++i) {
          const long long row = (long long)mirror_idx(sz + i, nz) * nx;
          for (int j = 0; j <= ORDER; ++j)
            acc += (double)pl[(row + mirror_idx(sx + j, nx)) * ec] * wz[i] * wx[j];
        }
        v = (R)acc;
      }
      pl[((long long)z * nx + x) * ec] = v;
    }
  }
}

// out[z][y][x] (+)= sum over the batch of slab[a][y][z][x] (float64 sum).
template <typename R, typename V>
__global__ void __launch_bounds__(256)
    k_sum_angles(const V *__restrict__ slab, int nb, V *__restrict__ out,
                 V *__restrict__ comp, int ny, int nz, int nx, int first,
                 int last) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long vol = (long long)ny * nz * nx;
  if (v >= vol) return;
  const int x = (int)(v % nx);
  const long long zy = v / nx;
  const int y = (int)(zy % ny), z = (int)(zy / ny);
  wide_t<V> sum;
  if (!first)
    sum = load_sum(out, comp, (size_t)v);
  else
    acc_zero(sum);
  const long long src = ((long long)y * nz + z) * nx + x;
  for (int a = 0; a < nb; ++a) sum = sum + widen(slab[a * vol + src]);
  store_sum(out, comp, (size_t)v, sum, !last);
}

// ---------------------------------------------------------------------------
// separable mirror-boundary B-spline prefilter, one thread per line
// (Unser/Thevenaz recursive filter; exact mirror initialisation), computed
// in float64.  Line l starts at (l / n0) * s1 + (l % n0) * s0, elements es.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_iir(R *data, long long nlines, int n, long long n0,
                      long long s0, long long s1, long long es, int npoles,
                      double z1, double z2, double gain) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines || n < 2) return;
  iir_line<R>(data + (l / n0) * s1 + (l % n0) * s0, n, es, npoles, z1, z2,
              gain);
}

// batched 1D transform of contiguous vectors
template <typename R>
__global__ void __launch_bounds__(256)
    k_fft_batch(const cpx<R> *in, cpx<R> *out, long long batch, int n,
                int logn, int dir, const cpx<R> *tw, int nb) {
  cpx<R> *buf = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *tmp = buf + nb * lds_len(n);
  const long long b0 = (long long)blockIdx.x * nb;
  for (int idx = threadIdx.x; idx < nb * n; idx += blockDim.x) {
    const long long b = b0 + idx / n;
    buf[idx] = b < batch ? in[b * n + idx % n] : mk<R>(0, 0);
  }
  __syncthreads();
  const cpx<R> *res = lds_transform<R>(buf, tmp, n, logn, nb, tw, dir);
  for (int idx = threadIdx.x; idx < nb * n; idx += blockDim.x) {
    const long long b = b0 + idx / n;
    if (b < batch) out[b * n + idx % n] = res[idx];
  }
}

// batched transform of contiguous vectors with the register radix-8 engine
template <typename R, int N, int DIR>
__global__ void __launch_bounds__(512)
    k_fft8_batch(const cpx<R> *in, cpx<R> *out, long long batch,
                 const cpx<R> *tw) {
  using FFT = Fft8<R, N, DIR>;
  constexpr int TT = FFT::T;
  const int tpb = blockDim.x / TT;
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *lds = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *b0 = lds + 2 * t * FFT::PADN, *b1 = b0 + FFT::PADN;
  const long long b = (long long)blockIdx.x * tpb + t;
  cpx<R> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
    v[m] = b < batch ? in[b * N + j + m * TT] : mk<R>(0, 0);
  FFT::run(v, b0, b1, tw, j);
  if (b < batch) {
#pragma unroll
    for (int m = 0; m < 8; ++m) out[b * N + j + m * TT] = v[m];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static size_t lds_bytes(int n, int nb, size_t esz) {
  return (size_t)2 * lds_len(n) * nb * esz;
}

static hipError_t set_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

template <typename R>
hipError_t launch_twiddles(cpx<R> *tw, int n, hipStream_t st) {
  k_twiddles<R><<<(n + 255) / 256, 256, 0, st>>>(tw, n);
  return hipGetLastError();
}

// Bluestein tables of length n (m = bluestein_len(n), tw already holding
// the m twiddles): the chirp w_j = exp(i pi j^2 / n) at tw + m, with j^2
// reduced mod 2n exactly, and the kernel h_d = conj(w_d) / m at positions d
// and m - d (zero between) at tw + m + n, transformed afterwards.
__device__ __forceinline__ cpx<double> chirp_of(long long j, int n) {
  double s, c;
  sincospi((double)((j * j) % (2LL * n)) / (double)n, &s, &c);
  return mk<double>(c, s);
}

__global__ void k_bluestein_chirp(cpx<double> *tw, int n, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  if (j < n) tw[m + j] = chirp_of(j, n);
  const int d = j < n ? j : (j > m - n ? m - j : -1);
  cpx<double> h = mk<double>(0, 0);
  if (d >= 0) {
    const cpx<double> w = chirp_of(d, n);
    h = mk<double>(w.re / m, -w.im / m);
  }
  tw[m + n + j] = h;
}

template <typename R>
__global__ void k_narrow_table(const cpx<double> *__restrict__ in, cpx<R> *__restrict__ out,
                               int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mk<R>((R)in[i].re, (R)in[i].im);
}

template <typename R>
hipError_t launch_fft_batch(const cpx<R> *in, cpx<R> *out, long long batch, int n,
                            int logn, int dir, const cpx<R> *tw, int nb,
                            hipStream_t st);

// Twiddle table of a transform of length n (dft_table_len(n) entries).
// Bluestein tables are built in float64 (in tw itself for float64 tables,
// else in scratch64, dft_table_len(n) entries) and rounded once.
template <typename R>
hipError_t launch_dft_table(cpx<R> *tw, int n, cpx<double> *scratch64, hipStream_t st) {
  if (!uses_bluestein(n)) return launch_twiddles<R>(tw, n, st);
  const int m = bluestein_len(n), len = dft_table_len(n);
  cpx<double> *s = sizeof(R) == sizeof(double) ? (cpx<double> *)tw : scratch64;
  if (!s) return hipErrorInvalidValue;
  hipError_t e = launch_twiddles<double>(s, m, st);
  if (e != hipSuccess) return e;
  k_bluestein_chirp<<<(m + 255) / 256, 256, 0, st>>>(s, n, m);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = launch_fft_batch<double>(s + m + n, s + m + n, 1, m, 31 - __builtin_clz(m), -1, s, 1,
                               st);
  if (e != hipSuccess) return e;
  if (sizeof(R) != sizeof(double)) {
    k_narrow_table<R><<<(len + 255) / 256, 256, 0, st>>>(s, tw, len);
    e = hipGetLastError();
  }
  return e;
}

template <typename RI, typename C, int N>
static hipError_t fwd_rows8_launch(const cpx<RI> *usin, const double *w, cpx<C> *F1,
                                   const cpx<C> *twx, const Geom &g, double pr, double pi,
                                   int a0, int na, hipStream_t st) {
  using FFT = Fft8<C, N, -1>;
  constexpr int TPB = 256 / FFT::T;
  const size_t sh = (size_t)TPB * FFT::NBUF * FFT::PADN * sizeof(cpx<C>);
  hipError_t e = set_lds((const void *)k_fwd_rows8<RI, C, N>, sh);
  if (e != hipSuccess) return e;
  const int rows = g.pad_mode == 1 ? g.Ny : g.lNy;  // edge: detector rows only
  dim3 grid((rows + TPB - 1) / TPB, na);
  k_fwd_rows8<RI, C, N><<<grid, 256, sh, st>>>(usin, w, F1, twx, g, pr, pi, a0);
  return hipGetLastError();
}

template <typename RI, typename C, int N>
static hipError_t fwd_cols8_launch(const cpx<C> *F1, cpx<RI> *S, const cpx<C> *twy,
                                   const cpx<C> *P, const Geom &g, int na, hipStream_t st) {
  using FFT = Fft8<C, N, -1>;
  constexpr int TPB = 256 / FFT::T;
  const size_t sh = (size_t)TPB * FFT::NBUF * FFT::PADN * sizeof(cpx<C>);
  hipError_t e = set_lds((const void *)k_fwd_cols8<RI, C, N>, sh);
  if (e != hipSuccess) return e;
  dim3 grid((g.Lx + TPB - 1) / TPB, na);
  k_fwd_cols8<RI, C, N><<<grid, 256, sh, st>>>(F1, S, twy, P, g);
  return hipGetLastError();
}

// ODTBRAIN_AMD_FWD_GENERIC: the forward stage keeps the LDS radix-2 engine
inline bool fwd_generic() {
  static const bool f = std::getenv("ODTBRAIN_AMD_FWD_GENERIC") != nullptr;
  return f;
}

template <typename RI, typename C>
hipError_t launch_fwd_rows(const cpx<RI> *usin, const double *w, cpx<C> *F1,
                           const cpx<C> *twx, const Geom &g, double pr,
                           double pi, int a0, int na, int nb,
                           hipStream_t st) {
  if (!fwd_generic()) switch (g.lNx) {
      case 64: return fwd_rows8_launch<RI, C, 64>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      case 128: return fwd_rows8_launch<RI, C, 128>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      case 256: return fwd_rows8_launch<RI, C, 256>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      case 512: return fwd_rows8_launch<RI, C, 512>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      case 1024: return fwd_rows8_launch<RI, C, 1024>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      case 2048: return fwd_rows8_launch<RI, C, 2048>(usin, w, F1, twx, g, pr, pi, a0, na, st);
      default: break;
    }
  const size_t sh = lds_bytes(g.lNx, nb, sizeof(cpx<C>));
  hipError_t e = set_lds((const void *)k_fwd_rows<RI, C>, sh);
  if (e != hipSuccess) return e;
  const int rows = g.pad_mode == 1 ? g.Ny : g.lNy;  // edge: detector rows only
  dim3 grid((rows + nb - 1) / nb, na);
  k_fwd_rows<RI, C><<<grid, 256, sh, st>>>(usin, w, F1, twx, g, pr, pi, a0, nb);
  return hipGetLastError();
}

template <typename RI, typename C>
hipError_t launch_fwd_cols(const cpx<C> *F1, cpx<RI> *S, const cpx<C> *twy,
                           const cpx<C> *P, const Geom &g, int na, int nb,
                           hipStream_t st) {
  if (!fwd_generic()) switch (g.lNy) {
      case 64: return fwd_cols8_launch<RI, C, 64>(F1, S, twy, P, g, na, st);
      case 128: return fwd_cols8_launch<RI, C, 128>(F1, S, twy, P, g, na, st);
      case 256: return fwd_cols8_launch<RI, C, 256>(F1, S, twy, P, g, na, st);
      case 512: return fwd_cols8_launch<RI, C, 512>(F1, S, twy, P, g, na, st);
      case 1024: return fwd_cols8_launch<RI, C, 1024>(F1, S, twy, P, g, na, st);
      case 2048: return fwd_cols8_launch<RI, C, 2048>(F1, S, twy, P, g, na, st);
      default: break;
    }
  const size_t sh = lds_bytes(g.lNy, nb, sizeof(cpx<C>));
  hipError_t e = set_lds((const void *)k_fwd_cols<RI, C>, sh);
  if (e != hipSuccess) return e;
  dim3 grid((g.Lx + nb - 1) / nb, na);
  k_fwd_cols<RI, C><<<grid, 256, sh, st>>>(F1, S, twy, P, g, nb);
  return hipGetLastError();
}

static constexpr size_t kLdsMax = 160 * 1024;

// support [j0, j0+len) of signed frequencies inside (-N/4, N/4)
static bool in_band(int j0, int len, int N) {
  return -j0 < N / 4 && j0 + len - 1 < N / 4;
}

// ODTBRAIN_AMD_COL_PER_ANGLE: column stage with one block per angle
// (k_col_fft8, propagator re-read per angle) instead of k_col_fft8_ang.
inline bool col_per_angle() {
  static const bool f = std::getenv("ODTBRAIN_AMD_COL_PER_ANGLE") != nullptr;
  return f;
}

template <typename R, int N>
static hipError_t col_fft8_launch(const cpx<R> *S, cpx<R> *T,
                                  const cpx<R> *twy, const double *pa,
                                  const PropAux<R> *aux, const R *q0z,
                                  const R *q0y, const Geom &g, int na, int ldt,
                                  const cpx<R> *gtab, hipStream_t st) {
  constexpr int TT = N / 8;
  // 256-thread blocks: at ~150 VGPRs a SIMD holds three waves, so three
  // 4-wave blocks fit per CU where only one 8-wave block would
  int tpb = 256 / TT < kColCpb ? 256 / TT : kColCpb;
  if (tpb < 1) tpb = 1;
  const size_t fft = ((size_t)Fft8<R, N, 1>::NBUF * tpb * lds_padded(N) + 4 * tpb) *
                     sizeof(cpx<R>);
  const size_t tile = (size_t)g.Ny * (kColCpb + 1) * sizeof(cpx<R>);
  const int staged = fft + tile <= 96 * 1024 ? 1 : 0;
  const size_t sh = fft + (staged ? tile : 0);
  if (sh > kLdsMax) return hipErrorInvalidValue;
  const bool band = in_band(g.jy0, g.Ly, N);
  // propagator in registers across the chunk's angles (see k_col_fft8_ang)
  if constexpr (TT <= 64) {
    const bool per_angle = g.col_mode == 1 || (g.col_mode == 0 && col_per_angle());
    if (band && gtab && staged && !per_angle) {
      constexpr int TPBa = 256 / TT < kColCpb ? 256 / TT : kColCpb;
      const size_t sha =
          ((size_t)TPBa * lds_padded(N) + (size_t)g.Ny * (kColCpb + 1)) * sizeof(cpx<R>);
      if (sha <= kLdsMax) {
        hipError_t e = set_lds((const void *)k_col_fft8_ang<R, N>, sha);
        if (e != hipSuccess) return e;
        dim3 grid((g.Lx + kColCpb - 1) / kColCpb, g.ln, 1);
        k_col_fft8_ang<R, N><<<grid, TPBa * TT, sha, st>>>(S, T, twy, q0y, g, na, ldt, gtab);
        return hipGetLastError();
      }
    }
  }
  // z-slices per block: measured at 256^2 x 360 (with the propagator
  // table): 1 -> 65.5 ms, 2 -> 64.6, 4 -> 64.6, 8 -> 67.8, 16 -> 67.4
  const int kb = 4;
  dim3 grid((g.Lx + kColCpb - 1) / kColCpb, (g.ln + kb - 1) / kb, na);
  const void *fn = band ? (const void *)k_col_fft8<R, N, true>
                        : (const void *)k_col_fft8<R, N, false>;
  hipError_t e = set_lds(fn, sh);
  if (e != hipSuccess) return e;
  if (band)
    k_col_fft8<R, N, true><<<grid, tpb * TT, sh, st>>>(S, T, twy, pa, aux, q0z,
                                                       q0y, g, kb, ldt, staged, gtab);
  else
    k_col_fft8<R, N, false><<<grid, tpb * TT, sh, st>>>(S, T, twy, pa, aux, q0z,
                                                        q0y, g, kb, ldt, staged, gtab);
  return hipGetLastError();
}

template <typename R, int N>
static size_t row_lds(int tpb) {
  return ((size_t)Fft8<R, N, 1>::NBUF * tpb * lds_padded(N) + 4 * tpb) * sizeof(cpx<R>);
}

// With one detector row (2D sinograms) the rows of T[a][k][0][.] and of the
// slab [a][0][k][.] are consecutive over the slices k, so the row stage
// views the slices as rows (Ny' = ln, ln' = 1): blocks then take many rows
// each instead of a single transform.
static Geom rows_view(const Geom &g) {
  Geom v = g;
  if (g.Ny == 1 && g.ln > 1) {
    v.Ny = g.ln;
    v.ln = 1;
  }
  return v;
}

template <typename R, typename V, int N>
static hipError_t row_fft8_launch(const cpx<R> *T, V *slab, const cpx<R> *twx,
                                  const R *q0x, const Geom &g0, int na,
                                  int slab0, int ldt, const cpx<R> *S2,
                                  const cpx<R> *gt, hipStream_t st) {
  const Geom g = rows_view(g0);
  constexpr int TT = N / 8;
  const int threads = TT < 256 ? 256 : TT;
  const int tpb = threads / TT;
  const size_t sh = row_lds<R, N>(tpb);
  if (sh > kLdsMax) return hipErrorInvalidValue;
  // 64 rows per block (16 iterations of 4 transforms for N = 512): the
  // prefetch of the next row runs through more of the block's work;
  // measured at 256^2 x 360: 8 rows 73.8 ms, 16 rows 69.8, 32 rows 68.9,
  // 64 rows 68.2 per reconstruction
  int rb = 64 < g.Ny ? 64 : g.Ny;
  dim3 grid((g.Ny + rb - 1) / rb, g.ln, na);
  const bool band = in_band(g.jx0, g.Lx, N);
  const void *fn = band ? (const void *)k_row_fft8<R, V, N, true>
                        : (const void *)k_row_fft8<R, V, N, false>;
  hipError_t e = set_lds(fn, sh);
  if (e != hipSuccess) return e;
  if (band)
    k_row_fft8<R, V, N, true><<<grid, threads, sh, st>>>(T, slab, twx, q0x, g,
                                                         rb, slab0, ldt, S2, gt);
  else
    k_row_fft8<R, V, N, false><<<grid, threads, sh, st>>>(T, slab, twx, q0x, g,
                                                          rb, slab0, ldt, S2, gt);
  return hipGetLastError();
}

// fast-path eligibility: power of two 64..4096 and the LDS footprint fits
template <typename R>
static bool fft8_col_ok(const Geom &g) {
  if (g.logNy < 6 || g.logNy > 12) return false;
  const int TT = g.lNy / 8;
  return (size_t)(TT > 64 ? 2 : 1) * lds_padded(g.lNy) * sizeof(cpx<R>) + 64 <= kLdsMax;
}
template <typename R>
static bool fft8_row_ok(const Geom &g) {
  if (g.logNx < 6 || g.logNx > 12) return false;
  const int TT = g.lNx / 8;
  return (size_t)(TT > 64 ? 2 : 1) * lds_padded(g.lNx) * sizeof(cpx<R>) + 64 <= kLdsMax;
}

template <typename R>
hipError_t launch_col_ifft(const cpx<R> *S, cpx<R> *T, const cpx<R> *twy,
                           const double *pa, const PropAux<R> *aux,
                           const R *q0z, const R *q0y, const Geom &g, int na,
                           int nb, int ldt, const cpx<R> *gtab,
                           hipStream_t st) {
  if (fft8_col_ok<R>(g)) {
    switch (g.lNy) {
      case 64: return col_fft8_launch<R, 64>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 128: return col_fft8_launch<R, 128>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 256: return col_fft8_launch<R, 256>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 512: return col_fft8_launch<R, 512>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 1024: return col_fft8_launch<R, 1024>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 2048: return col_fft8_launch<R, 2048>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
      case 4096: return col_fft8_launch<R, 4096>(S, T, twy, pa, aux, q0z, q0y, g, na, ldt, gtab, st);
    }
  }
  if (g.lNy == 1 && g.Ny == 1 && g.Ly == 1 && !g.fuse_y) {
    constexpr int kb = 16;
    dim3 grid((g.Lx + 255) / 256, (g.ln + kb - 1) / kb, na);
    k_prop_2d<R><<<grid, 256, 0, st>>>(S, T, pa, aux, q0z, g, kb, ldt, gtab);
    return hipGetLastError();
  }
  const size_t sh = lds_bytes(g.lNy, nb, sizeof(cpx<R>));
  hipError_t e = set_lds((const void *)k_col_ifft<R>, sh);
  if (e != hipSuccess) return e;
  dim3 grid(g.Lx, (g.ln + nb - 1) / nb, na);
  k_col_ifft<R><<<grid, 256, sh, st>>>(S, T, twy, pa, aux, q0z, q0y, g, nb, ldt, gtab);
  return hipGetLastError();
}

// 2D plans whose rows fuse k_prop_2d's product (S2, gt non-null in
// launch_row_ifft): the register engine and the propagator table.
template <typename R>
bool row_fuses_2d(const Geom &g, bool gtab) {
  return gtab && g.lNy == 1 && g.Ny == 1 && g.Ly == 1 && !g.fuse_y && fft8_row_ok<R>(g) &&
         g.lNx <= 4096;
}

template <typename R, typename V>
hipError_t launch_row_ifft(const cpx<R> *T, V *slab, const cpx<R> *twx,
                           const R *q0x, const Geom &g, int na, int nb,
                           int slab0, int ldt, hipStream_t st, const cpx<R> *S2,
                           const cpx<R> *gt) {
  if ((S2 || gt) && !(S2 && gt && row_fuses_2d<R>(g, true))) return hipErrorInvalidValue;
  if (fft8_row_ok<R>(g)) {
    switch (g.lNx) {
      case 64: return row_fft8_launch<R, V, 64>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 128: return row_fft8_launch<R, V, 128>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 256: return row_fft8_launch<R, V, 256>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 512: return row_fft8_launch<R, V, 512>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 1024: return row_fft8_launch<R, V, 1024>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 2048: return row_fft8_launch<R, V, 2048>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
      case 4096: return row_fft8_launch<R, V, 4096>(T, slab, twx, q0x, g, na, slab0, ldt, S2, gt, st);
    }
  }
  const Geom gv = rows_view(g);
  const size_t sh = lds_bytes(gv.lNx, nb, sizeof(cpx<R>));
  hipError_t e = set_lds((const void *)k_row_ifft<R, V>, sh);
  if (e != hipSuccess) return e;
  dim3 grid((gv.Ny + nb - 1) / nb, gv.ln, na);
  k_row_ifft<R, V><<<grid, 256, sh, st>>>(T, slab, twx, q0x, gv, nb, slab0, ldt);
  return hipGetLastError();
}

size_t rotate_window_entries(int nz, int nx) {
  return (size_t)((nz + kRotTile - 1) / kRotTile) * ((nx + kRotTile - 1) / kRotTile);
}

// Planes per block: 4 when both tile buffers then fit in 32 KiB, else 2.
// At 32 KiB per block a CU holds five blocks (20 waves); measured at
// 256^2 x 360 (complex64): 2 planes 21.1 ms, 3 planes 21.7 ms, 4 planes
// (64 KiB, two blocks per CU) 22.5 ms, 1 plane 26.3 ms per reconstruction.
template <typename V>
constexpr int rot_yc_max() {
  return 2 * 4 * kRotWin * kRotWin * (int)sizeof(V) <= 32768 ? 4 : 2;
}

// Complex samples with at least four planes take k_rotate_lds2 (four
// planes, pipelined half buffers); measured at 256^2 x 360 (complex64,
// 48 angles per launch) 2.02 ms per launch against 2.10 ms for
// k_rotate_lds with two planes.  ODTBRAIN_AMD_ROTATE_TWO_PLANE forces
// k_rotate_lds (for comparisons).
inline bool rotate_two_plane() {
  static const bool f = std::getenv("ODTBRAIN_AMD_ROTATE_TWO_PLANE") != nullptr;
  return f;
}

// groups > 1 (with part: room for 2 * groups volumes): the angles are split
// into that many contiguous groups that run side by side (grid y), each
// into its own compensated sum in part, and k_fold_groups adds them to out
// in group order.  For planes with few tiles (2D images), whose blocks
// would otherwise each loop over every angle on a fraction of the CUs.
template <typename R, typename V, int ORDER, int YC, bool PIPE = false>
static hipError_t rotate_lds_launch(const V *coef, int nb, const double *params,
                                    int *win, V *out, V *comp, int ny, int nz,
                                    int nx, int first, int last, hipStream_t st,
                                    int groups = 1, V *part = nullptr) {
  static_assert(!PIPE || YC == 4, "the pipelined kernel holds four planes");
  if (nb < 1) return hipSuccess;
  const int ntx = (nx + kRotTile - 1) / kRotTile;
  const int ntz = (nz + kRotTile - 1) / kRotTile;
  const int nt = ntx * ntz;
  const long long nw = (long long)nb * nt;
  const int nyc = (ny + YC - 1) / YC;
  const int xcd_map = nyc >= 8 ? 1 : 0;
  const long long blocks =
      xcd_map ? 8LL * nt * ((nyc + 7) / 8) : (long long)nt * nyc;
  // 16-byte loads need windows inside the plane and 16-byte aligned rows
  constexpr int AL = 16 / (int)sizeof(V);
  const bool vec4 = nz >= kRotWin && nx >= kRotWin && nx % AL == 0 &&
                    ((uintptr_t)coef & 15) == 0;
  k_rot_windows<ORDER><<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(
      params, nb, ntz, ntx, nz, nx, vec4 ? AL : 1, (int4 *)win);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if constexpr (PIPE) {
    if (vec4)
      k_rotate_lds2<R, V, ORDER, 4><<<(unsigned)blocks, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, out, comp, ny, nz, nx, first, last,
          ntx, ntz, xcd_map);
    else
      k_rotate_lds2<R, V, ORDER, 1><<<(unsigned)blocks, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, out, comp, ny, nz, nx, first, last,
          ntx, ntz, xcd_map);
  } else if (groups > 1 && part && nb > 1) {
    const int ag = (nb + groups - 1) / groups;
    const int G = (nb + ag - 1) / ag;  // every group non-empty
    const size_t n = (size_t)ny * nz * nx;
    const dim3 grid((unsigned)blocks, (unsigned)G);
    if (vec4)
      k_rotate_lds<R, V, ORDER, YC, 4><<<grid, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, part, part + (size_t)G * n, ny, nz,
          nx, 1, 0, ntx, ntz, xcd_map, ag, n);
    else
      k_rotate_lds<R, V, ORDER, YC, 1><<<grid, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, part, part + (size_t)G * n, ny, nz,
          nx, 1, 0, ntx, ntz, xcd_map, ag, n);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_fold_groups<V><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        out, comp, part, G, n, first, last);
  } else {
    if (vec4)
      k_rotate_lds<R, V, ORDER, YC, 4><<<(unsigned)blocks, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, out, comp, ny, nz, nx, first, last,
          ntx, ntz, xcd_map, 0, 0);
    else
      k_rotate_lds<R, V, ORDER, YC, 1><<<(unsigned)blocks, 256, 0, st>>>(
          coef, nb, params, (const int4 *)win, out, comp, ny, nz, nx, first, last,
          ntx, ntz, xcd_map, 0, 0);
  }
  return hipGetLastError();
}

template <typename R, typename V, int ORDER>
static hipError_t rotate_dispatch_yc(const V *coef, int nb, const double *params,
                                     int *win, V *out, V *comp, int ny, int nz,
                                     int nx, int first, int last, hipStream_t st,
                                     int groups, V *part) {
  if constexpr (std::is_same<V, cpx<R>>::value) {
    if (ny >= 4 && !rotate_two_plane())
      return rotate_lds_launch<R, V, ORDER, 4, true>(coef, nb, params, win, out,
                                                     comp, ny, nz, nx, first,
                                                     last, st);
  }
  constexpr int YM = rot_yc_max<V>();
  if (ny >= YM)
    return rotate_lds_launch<R, V, ORDER, YM>(coef, nb, params, win, out, comp,
                                              ny, nz, nx, first, last, st);
  if (ny >= 2)
    return rotate_lds_launch<R, V, ORDER, 2>(coef, nb, params, win, out, comp,
                                             ny, nz, nx, first, last, st);
  return rotate_lds_launch<R, V, ORDER, 1>(coef, nb, params, win, out, comp, ny,
                                           nz, nx, first, last, st, groups, part);
}

template <typename R, typename V>
hipError_t launch_rotate_acc(int order, const V *coef, int nb,
                             const double *params, int *win, V *out, V *comp,
                             int ny, int nz, int nx, int first, int last,
                             hipStream_t st, int groups, V *part) {
  switch (order) {
    case 0: return rotate_dispatch_yc<R, V, 0>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
    case 1: return rotate_dispatch_yc<R, V, 1>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
    case 2: return rotate_dispatch_yc<R, V, 2>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
    case 3: return rotate_dispatch_yc<R, V, 3>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
    case 4: return rotate_dispatch_yc<R, V, 4>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
    case 5: return rotate_dispatch_yc<R, V, 5>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st, groups, part);
  }
  return hipErrorInvalidValue;
}

// ODTBRAIN_AMD_AFFINE_DIRECT: the tilted gather from global memory
// (k_affine_acc) instead of the LDS-staged k_affine_lds.
inline bool affine_direct() {
  static const bool f = std::getenv("ODTBRAIN_AMD_AFFINE_DIRECT") != nullptr;
  return f;
}

template <typename R, typename V, int ORDER>
static hipError_t affine_dispatch(const V *coef, int nb, const double *params,
                                  int *win, V *out, V *comp, int ny, int nz, int nx,
                                  int first, int last, hipStream_t st) {
  // one reflection is the whole mirror rule when every edge is >= 4
  // LDS staging for samples of at most 8 bytes (complex64, real types): the
  // complex128 window and its registers leave too few blocks per CU
  // (measured at 192^2 x 240: FP64 84 -> 191 ms per reconstruction)
  if (sizeof(V) <= 8 && win && nx >= 4 && ny >= 4 && nz >= 4 && nb >= 1 &&
      !affine_direct()) {
    const int ntx = (nx + kAffTile - 1) / kAffTile, nty = (ny + kAffTile - 1) / kAffTile,
              ntz = (nz + kAffTile - 1) / kAffTile;
    const long long nt = (long long)ntx * nty * ntz;
    const long long nw = nt * nb;
    k_affine_windows<ORDER><<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(
        params, nb, ntx, nty, ntz, ny, nz, nx, (int4 *)win);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int xcd_map = nt % 8 == 0 ? 1 : 0;
    k_affine_lds<R, V, ORDER><<<(unsigned)nt, 256, 0, st>>>(
        coef, nb, params, (const int4 *)win, out, comp, ny, nz, nx, first, last, ntx, nty,
        ntz, xcd_map);
    return hipGetLastError();
  }
  dim3 block(256);
  dim3 grid((nx + 7) / 8, (ny + 3) / 4, (nz + 7) / 8);
  k_affine_acc<R, V, ORDER><<<grid, block, 0, st>>>(coef, nb, params, out, comp,
                                                    ny, nz, nx, first, last);
  return hipGetLastError();
}

template <typename R, typename V>
hipError_t launch_affine_acc(int order, const V *coef, int nb,
                             const double *params, int *win, V *out, V *comp, int ny,
                             int nz, int nx, int first, int last,
                             hipStream_t st) {
  switch (order) {
    case 0: return affine_dispatch<R, V, 0>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
    case 1: return affine_dispatch<R, V, 1>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
    case 2: return affine_dispatch<R, V, 2>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
    case 3: return affine_dispatch<R, V, 3>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
    case 4: return affine_dispatch<R, V, 4>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
    case 5: return affine_dispatch<R, V, 5>(coef, nb, params, win, out, comp, ny, nz, nx, first, last, st);
  }
  return hipErrorInvalidValue;
}

template <typename R, typename V>
hipError_t launch_seq_rotate_sum(int order, V *slab, int nb,
                                 const double *params, V *out, V *comp,
                                 int ny, int nz, int nx, int first, int last,
                                 hipStream_t st) {
  const int ec = sizeof(V) / sizeof(R);
  const long long njobs = (long long)nb * ny * ec;
  const unsigned blocks = (unsigned)((njobs + 63) / 64);
  if (order == 0)
    k_seq_rotate<R, 0><<<blocks, 64, 0, st>>>((R *)slab, njobs, ny, nz, nx, ec,
                                              params);
  else if (order == 1)
    k_seq_rotate<R, 1><<<blocks, 64, 0, st>>>((R *)slab, njobs, ny, nz, nx, ec,
                                              params);
  else
    return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long vol = (long long)ny * nz * nx;
  k_sum_angles<R, V><<<(unsigned)((vol + 255) / 256), 256, 0, st>>>(
      slab, nb, out, comp, ny, nz, nx, first, last);
  return hipGetLastError();
}

bool iir_poles(int order, int *np, double *z1, double *z2, double *gain) {
  *np = 1;
  *z2 = 0;
  switch (order) {
    case 2: *z1 = sqrt(8.0) - 3.0; break;
    case 3: *z1 = sqrt(3.0) - 2.0; break;
    case 4:
      *np = 2;
      *z1 = sqrt(664.0 - sqrt(438976.0)) + sqrt(304.0) - 19.0;
      *z2 = sqrt(664.0 + sqrt(438976.0)) - sqrt(304.0) - 19.0;
      break;
    case 5:
      *np = 2;
      *z1 = sqrt(67.5 - sqrt(4436.25)) + sqrt(26.25) - 6.5;
      *z2 = sqrt(67.5 + sqrt(4436.25)) - sqrt(26.25) - 6.5;
      break;
    default: return false;
  }
  *gain = (1.0 - *z1) * (1.0 - 1.0 / *z1);
  if (*np == 2) *gain *= (1.0 - *z2) * (1.0 - 1.0 / *z2);
  return true;
}

template <typename R>
hipError_t launch_iir(R *data, long long nlines, int n, long long n0,
                      long long s0, long long s1, long long es, int order,
                      hipStream_t st) {
  double z1 = 0, z2 = 0, gain = 1;
  int np = 0;
  if (!iir_poles(order, &np, &z1, &z2, &gain)) return hipErrorInvalidValue;
  if (nlines <= 0) return hipSuccess;
  const long long blocks = (nlines + 255) / 256;
  k_iir<R><<<(unsigned)blocks, 256, 0, st>>>(data, nlines, n, n0, s0, s1, es,
                                            np, z1, z2, gain);
  return hipGetLastError();
}

template <typename R, int N>
static hipError_t fft8_batch_launch(const cpx<R> *in, cpx<R> *out,
                                    long long batch, int dir, const cpx<R> *tw,
                                    hipStream_t st) {
  constexpr int TT = N / 8;
  const int threads = TT < 256 ? 256 : TT;
  const int tpb = threads / TT;
  const size_t sh = 2 * (size_t)tpb * lds_padded(N) * sizeof(cpx<R>);
  if (sh > kLdsMax) return hipErrorInvalidValue;
  const void *fn = dir < 0 ? (const void *)k_fft8_batch<R, N, -1>
                           : (const void *)k_fft8_batch<R, N, +1>;
  hipError_t e = set_lds(fn, sh);
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)((batch + tpb - 1) / tpb);
  if (dir < 0)
    k_fft8_batch<R, N, -1><<<blocks, threads, sh, st>>>(in, out, batch, tw);
  else
    k_fft8_batch<R, N, +1><<<blocks, threads, sh, st>>>(in, out, batch, tw);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_fft_batch(const cpx<R> *in, cpx<R> *out, long long batch,
                            int n, int logn, int dir, const cpx<R> *tw, int nb,
                            hipStream_t st) {
  if (logn >= 3 && logn <= 12 &&
      2 * (size_t)(n / 8 < 256 ? 256 / (n / 8) : 1) * lds_padded(n) *
              sizeof(cpx<R>) <= kLdsMax) {
    switch (n) {
      case 8: return fft8_batch_launch<R, 8>(in, out, batch, dir, tw, st);
      case 16: return fft8_batch_launch<R, 16>(in, out, batch, dir, tw, st);
      case 32: return fft8_batch_launch<R, 32>(in, out, batch, dir, tw, st);
      case 64: return fft8_batch_launch<R, 64>(in, out, batch, dir, tw, st);
      case 128: return fft8_batch_launch<R, 128>(in, out, batch, dir, tw, st);
      case 256: return fft8_batch_launch<R, 256>(in, out, batch, dir, tw, st);
      case 512: return fft8_batch_launch<R, 512>(in, out, batch, dir, tw, st);
      case 1024: return fft8_batch_launch<R, 1024>(in, out, batch, dir, tw, st);
      case 2048: return fft8_batch_launch<R, 2048>(in, out, batch, dir, tw, st);
      case 4096: return fft8_batch_launch<R, 4096>(in, out, batch, dir, tw, st);
    }
  }
  const size_t sh = lds_bytes(n, nb, sizeof(cpx<R>));
  hipError_t e = set_lds((const void *)k_fft_batch<R>, sh);
  if (e != hipSuccess) return e;
  const long long blocks = (batch + nb - 1) / nb;
  k_fft_batch<R><<<(unsigned)blocks, 256, sh, st>>>(in, out, batch, n, logn,
                                                   dir, tw, nb);
  return hipGetLastError();
}

// explicit instantiations
template hipError_t launch_fwd_rows<float, double>(
    const cpx<float> *, const double *, cpx<double> *, const cpx<double> *,
    const Geom &, double, double, int, int, int, hipStream_t);
template hipError_t launch_fwd_cols<float, double>(
    const cpx<double> *, cpx<float> *, const cpx<double> *, const cpx<double> *,
    const Geom &, int, int, hipStream_t);
#define ODT_INST(R)                                                            \
  template hipError_t launch_twiddles<R>(cpx<R> *, int, hipStream_t);          \
  template hipError_t launch_dft_table<R>(cpx<R> *, int, cpx<double> *,       \
                                          hipStream_t);                      \
  template hipError_t launch_fwd_rows<R, R>(const cpx<R> *, const double *,    \
                                            cpx<R> *, const cpx<R> *,          \
                                            const Geom &, double, double, int, \
                                            int, int, hipStream_t);            \
  template hipError_t launch_fwd_cols<R, R>(const cpx<R> *, cpx<R> *,          \
                                            const cpx<R> *, const cpx<R> *,    \
                                            const Geom &, int, int,            \
                                            hipStream_t);                      \
  template hipError_t launch_col_ifft<R>(                                      \
      const cpx<R> *, cpx<R> *, const cpx<R> *, const double *,                \
      const PropAux<R> *, const R *, const R *, const Geom &, int, int, int,   \
      const cpx<R> *, hipStream_t);                                            \
  template hipError_t launch_prop_table<R>(const double *, const PropAux<R> *, \
                                           const R *, const Geom &, cpx<R> *,  \
                                           hipStream_t);                                                            \
  template hipError_t launch_row_ifft<R, R>(const cpx<R> *, R *,               \
                                            const cpx<R> *, const R *,         \
                                            const Geom &, int, int, int, int,  \
                                            hipStream_t, const cpx<R> *,       \
                                            const cpx<R> *);                   \
  template hipError_t launch_row_ifft<R, cpx<R>>(                              \
      const cpx<R> *, cpx<R> *, const cpx<R> *, const R *, const Geom &, int,  \
      int, int, int, hipStream_t, const cpx<R> *, const cpx<R> *);             \
  template bool row_fuses_2d<R>(const Geom &, bool);                           \
  template hipError_t launch_rotate_acc<R, R>(int, const R *, int,             \
                                              const double *, int *, R *, R *, \
                                              int, int, int, int, int,         \
                                              hipStream_t, int, R *);          \
  template hipError_t launch_rotate_acc<R, cpx<R>>(                            \
      int, const cpx<R> *, int, const double *, int *, cpx<R> *, cpx<R> *,     \
      int, int, int, int, int, hipStream_t, int, cpx<R> *);                    \
  template hipError_t launch_affine_acc<R, R>(int, const R *, int,             \
                                              const double *, int *, R *, R *, \
                                              int, int, int, int, int,         \
                                              hipStream_t);                    \
  template hipError_t launch_affine_acc<R, cpx<R>>(                            \
      int, const cpx<R> *, int, const double *, int *, cpx<R> *, cpx<R> *,     \
      int, int, int, int, int, hipStream_t);                                   \
  template hipError_t launch_seq_rotate_sum<R, R>(int, R *, int,              \
                                                  const double *, R *, R *,    \
                                                  int, int, int, int, int,     \
                                                  hipStream_t);                \
  template hipError_t launch_seq_rotate_sum<R, cpx<R>>(                        \
      int, cpx<R> *, int, const double *, cpx<R> *, cpx<R> *, int, int, int,   \
      int, int, hipStream_t);                                                  \
  template hipError_t launch_iir<R>(R *, long long, int, long long, long long, \
                                    long long, long long, int, hipStream_t);   \
  template hipError_t launch_fft_batch<R>(const cpx<R> *, cpx<R> *, long long, \
                                          int, int, int, const cpx<R> *, int,  \
                                          hipStream_t);

ODT_INST(float)
ODT_INST(double)

}  // namespace odt

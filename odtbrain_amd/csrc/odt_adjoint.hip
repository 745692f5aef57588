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
//   k_adj_rotate     R_a^T g: the rotate gather transposed (same
//                    coordinates, weights and mirrored taps; no atomics)
//   k_adj_fir +      Q^T = D Q D^{-1} per axis, D = diag(1/2, 1, ..., 1, 1/2)
//   k_adj_fir_rows   (whole-sample mirror extension; checked numerically);
//                    plans with the fused prefilter (orders 2, 3) skip them:
//                    the next two kernels carry the fused operator's transpose
//   k_adj_row        forward transform along x of each zero-padded slab row
//                    (plus the fused x-prefilter's transposed boundary
//                    terms), support columns kept (IDFT_x^H = DFT_x)
//   k_adj_col        forward transform along y of each support column and
//                    each z, times the conjugate column-stage propagator
//                    (the plan's table: e^{i k_z z}, fused z-prefilter terms
//                    included), summed over z in registers
//   k_adj_inv_col    times conj(P), inverse transform along y, Pad^T in y
//   k_adj_inv_row    inverse transform along x, Pad^T in x, times w_a
// Transforms use the register radix-8 engine for lengths 64..512 (one
// transform per wavefront or less).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "odt_device.h"
#include "odt_fft8.h"
#include "odt_launch.h"

namespace odt {

extern __shared__ __align__(16) unsigned char odt_smem[];

// R_a^T as a gather: the coefficient (a, y, jz, jx) sums w * g over the
// voxels whose forward taps (k_rotate_lds / scipy's rotate: same float64
// coordinates, strict bounds test, tap start and mirrored taps) include it,
// directly or through the whole-sample mirror (virtual taps -j and
// 2n - 2 - j).  Those voxels lie in the inverse-rotated square of half
// width (ORDER + 1) / 2 about the tap; the candidates of its bounding box
// are tested exactly.  The rotation is about y, so one candidate search
// serves YC planes.  Block: a 16 x 16 coefficient tile of one angle and
// YC planes; every coefficient is written (no atomics, no clearing).
template <typename R, int ORDER, int YC>
__global__ void __launch_bounds__(256)
    k_adj_rotate(const cpx<R> *__restrict__ g, const double *__restrict__ params,
                 cpx<R> *__restrict__ slab, int ny, int nz, int nx) {
  const int ntx = (nx + 15) >> 4;
  const int tz = blockIdx.x / ntx, tx = blockIdx.x - tz * ntx;
  const int jz = tz * 16 + (threadIdx.x >> 4);
  const int jx = tx * 16 + (threadIdx.x & 15);
  if (jz >= nz || jx >= nx) return;
  const int a = blockIdx.y, y0 = blockIdx.z * YC;
  const int yc = min(YC, ny - y0);
  const double *p = params + 6 * (size_t)a;
  const double m00 = p[0], m01 = p[1], m10 = p[2], m11 = p[3];
  const double oz = p[4], ox = p[5];
  constexpr double h = 0.5 * (ORDER + 1);
  constexpr int K = ORDER / 2 + 1;  // taps reach at most K past an edge
  const double hz = h * (fabs(m00) + fabs(m10)) + 1e-7;
  const double hx = h * (fabs(m01) + fabs(m11)) + 1e-7;
  const int nzc = (int)floor(2.0 * hz) + 1, nxc = (int)floor(2.0 * hx) + 1;
  constexpr int NONE = -(1 << 30);
  const int vz1 = (jz >= 1 && jz <= K)                    ? -jz
                  : (jz <= nz - 2 && jz >= nz - 1 - K) ? 2 * nz - 2 - jz
                                                        : NONE;
  const int vx1 = (jx >= 1 && jx <= K)                    ? -jx
                  : (jx <= nx - 2 && jx >= nx - 1 - K) ? 2 * nx - 2 - jx
                                                        : NONE;
  cpx<R> acc[YC];
#pragma unroll
  for (int k = 0; k < YC; ++k) acc[k] = mk<R>(0, 0);
  for (int iv = 0; iv < 4; ++iv) {
    const int vz = (iv & 1) ? vz1 : jz, vx = (iv & 2) ? vx1 : jx;
    if (vz == NONE || vx == NONE) continue;
    const double dz = (double)vz - oz, dx = (double)vx - ox;
    const int zlo = (int)ceil(m00 * dz + m10 * dx - hz);
    const int xlo = (int)ceil(m01 * dz + m11 * dx - hx);
    for (int iz = 0; iz < nzc; ++iz) {
      const int z = zlo + iz;
      if (z < 0 || z >= nz) continue;
      for (int ix = 0; ix < nxc; ++ix) {
        const int x = xlo + ix;
        if (x < 0 || x >= nx) continue;
        const double cz = affine2((double)z, (double)x, m00, m01, oz);
        const double cx = affine2((double)z, (double)x, m10, m11, ox);
        if (cz < 0.0 || cz > (double)(nz - 1) || cx < 0.0 ||
            cx > (double)(nx - 1))
          continue;
        const unsigned tzi = (unsigned)(vz - spline_start<ORDER>(cz));
        const unsigned txi = (unsigned)(vx - spline_start<ORDER>(cx));
        if (tzi > (unsigned)ORDER || txi > (unsigned)ORDER) continue;
        const R w = (R)(bspline_val<ORDER>(cz - (double)vz) *
                        bspline_val<ORDER>(cx - (double)vx));
        const cpx<R> *gp = g + ((size_t)z * ny + y0) * nx + x;
#pragma unroll
        for (int k = 0; k < YC; ++k)
          if (k < yc) acc[k] = acc[k] + w * gp[(size_t)k * nx];
      }
    }
  }
  cpx<R> *dst = slab + (((size_t)a * ny + y0) * nz + jz) * nx + jx;
#pragma unroll
  for (int k = 0; k < YC; ++k)
    if (k < yc) dst[(size_t)k * nz * nx] = acc[k];
}

// Q^T = D Q D^{-1} along one axis of complex lines, out of place, with Q
// the mirror-boundary B-spline prefilter written as a convolution: the
// recursive filter with its whole-line mirror initialisation equals the
// two-sided impulse response h of the filter applied to the infinitely
// mirrored line (h decays as z^|k|; the host truncates it where it falls
// below 1e-18 of h[0]).  Line l starts at (l / n0) * s1 + (l % n0) * s0,
// element i at + i * es (complex units); consecutive threads take
// consecutive lines, each thread kFirOB consecutive outputs, streaming the
// kFirOB + 2K inputs once.  hz holds h[k] at k + K + kFirOB, zero-padded
// to 2K + 2 kFirOB entries, so every (output, input) pair indexes inside;
// its index is uniform, so the taps come through the scalar cache.
constexpr int kFirOB = 8;
constexpr int adj_fir_pad_c = kFirOB;

template <typename R>
__global__ void __launch_bounds__(256)
    k_adj_fir(const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out,
              long long nlines, int n, long long n0, long long s0, long long s1,
              long long es, const R *__restrict__ hz, int K, int nzb) {
  // Blocks are dealt round-robin to the 8 XCDs: all nzb output blocks of a
  // group of 256 lines go to one XCD, so their overlapping input windows
  // meet in its L2.
  const int bid = blockIdx.x, xcd = bid & 7, k = bid >> 3;
  const long long grp = (long long)(k / nzb) * 8 + xcd;
  const int i0 = (k % nzb) * kFirOB;
  const long long l = grp * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  const long long base = (l / n0) * s1 + (l % n0) * s0;
  const cpx<R> *src = in + base;
  cpx<R> acc[kFirOB];
#pragma unroll
  for (int o = 0; o < kFirOB; ++o) acc[o] = mk<R>(0, 0);
  const int jn = kFirOB + 2 * K;
  // inputs in groups of 8 loads issued together, then their updates
  for (int j0 = 0; j0 < jn; j0 += 8) {
    cpx<R> v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      v[u] = mk<R>(0, 0);
      if (j0 + u < jn) {
        const int m = mirror_idx(i0 - K + j0 + u, n);
        v[u] = src[(long long)m * es];
        if (m == 0 || m == n - 1) v[u] = mk<R>(R(2) * v[u].re, R(2) * v[u].im);  // D^-1
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (j0 + u < jn) {
        const R *w = hz + (2 * K + kFirOB - (j0 + u));
#pragma unroll
        for (int o = 0; o < kFirOB; ++o) acc[o] = acc[o] + w[o] * v[u];
      }
    }
  }
#pragma unroll
  for (int o = 0; o < kFirOB; ++o) {
    const int i = i0 + o;
    if (i < n) {
      cpx<R> r = acc[o];
      if (i == 0 || i == n - 1) r = mk<R>(R(0.5) * r.re, R(0.5) * r.im);  // D
      out[base + (long long)i * es] = r;
    }
  }
}

// The same convolution along contiguous rows of length n: each wave
// copies one row, mirror-extended by K on both sides and scaled by D^-1 at
// its ends, into LDS and its lanes produce consecutive outputs from there.
template <typename R>
__global__ void __launch_bounds__(256)
    k_adj_fir_rows(const cpx<R> *__restrict__ in, cpx<R> *__restrict__ out,
                   long long nrows, int n, const R *__restrict__ hz, int K) {
  cpx<R> *ext = reinterpret_cast<cpx<R> *>(odt_smem);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ne = n + 2 * K;
  cpx<R> *e = ext + (size_t)wave * ne;
  const long long row = (long long)blockIdx.x * 4 + wave;
  const bool live = row < nrows;
  if (live) {
    const cpx<R> *src = in + row * n;
    for (int i = lane; i < ne; i += 64) {
      const int m = mirror_idx(i - K, n);
      cpx<R> v = src[m];
      if (m == 0 || m == n - 1) v = mk<R>(R(2) * v.re, R(2) * v.im);  // D^-1
      e[i] = v;
    }
  }
  __syncthreads();
  if (!live) return;
  const R *h = hz + adj_fir_pad_c;  // h[k] at h[k + K]
  for (int x = lane; x < n; x += 64) {
    cpx<R> acc = mk<R>(0, 0);
    const cpx<R> *ex = e + x + 2 * K;  // ex[-k - K] = ext[x + K - k]
#pragma unroll 8
    for (int k = 0; k <= 2 * K; ++k) acc = acc + h[k] * ex[-k];
    if (x == 0 || x == n - 1) acc = mk<R>(R(0.5) * acc.re, R(0.5) * acc.im);  // D
    out[row * n + x] = acc;
  }
}

// Q^T = D Q D^{-1} along x (D = diag(1/2, 1, ..., 1, 1/2), Q the mirror
// B-spline prefilter of k_iir) for contiguous complex rows of length n:
// a block stages rb rows in LDS (row stride 2n + 2 values, so the 2 rb
// lines of a wave fall in distinct banks) and runs one line per lane.
template <typename R>
__global__ void __launch_bounds__(64)
    k_adj_iir_rows(R *__restrict__ data, long long nrows, int n, int rb,
                   int npoles, double z1, double z2, double gain) {
  R *buf = reinterpret_cast<R *>(odt_smem);
  const int ld = 2 * n + 2, w = 2 * n;
  const long long r0 = (long long)blockIdx.x * rb;
  const int nr = (int)min((long long)rb, nrows - r0);
  R *src = data + r0 * w;
  for (int i = threadIdx.x; i < nr * w; i += 64) {
    const int r = i / w;
    buf[r * ld + (i - r * w)] = src[i];
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * nr && n >= 2) {
    R *q = buf + (threadIdx.x >> 1) * ld + (threadIdx.x & 1);
    q[0] *= R(2);
    q[2 * (n - 1)] *= R(2);
    iir_line<R>(q, n, 2, npoles, z1, z2, gain);
    q[0] *= R(0.5);
    q[2 * (n - 1)] *= R(0.5);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr * w; i += 64) {
    const int r = i / w;
    src[i] = buf[r * ld + (i - r * w)];
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
// Fused x-prefilter plans (g.fuse_x): the forward row stage crops the
// inverse transform h and adds q0x[x] R0 + q0x[nx-1-x] R1 with
// R0 = b1 (h[padxl-1] - h[padxl+1]), R1 = b1 (h[padxl+nx] - h[padxl+nx-2]);
// its transpose adds b1 <q0x, row> at padxl-1 (minus at padxl+1) and
// b1 <reversed q0x, row> at padxl+nx (minus at padxl+nx-2) to the
// zero-padded row before the transform (q0x and b1 are real).
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_row(const cpx<R> *__restrict__ slab, cpx<R> *__restrict__ Tp,
              const cpx<R> *__restrict__ twx, Geom g, int rb, int ldt,
              const R *__restrict__ q0x) {
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
  // fused x-prefilter: the lane's q0x weights and, per slot, the signed
  // factor (+-b1 or 0) of the two boundary sums it receives (the same for
  // every row)
  R qa[8], qb[8], fa[8], fb[8];
  const bool fx = g.fuse_x;
  if (fx) {
    const R b1 = (R)g.b1;
    const int p0 = wrap(g.padxl - 1, N), p1 = wrap(g.padxl + 1, N);
    const int p2 = wrap(g.padxl + nx, N), p3 = wrap(g.padxl + nx - 2, N);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int x = j + m * TT - g.padxl, pm = j + m * TT;
      const bool in = x >= 0 && x < nx;
      qa[m] = in ? q0x[x] : R(0);
      qb[m] = in ? q0x[nx - 1 - x] : R(0);
      fa[m] = (pm == p0 ? b1 : R(0)) - (pm == p1 ? b1 : R(0));
      fb[m] = (pm == p2 ? b1 : R(0)) - (pm == p3 ? b1 : R(0));
    }
  }
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
    if (fx) {
      cpx<R> s0 = mk<R>(0, 0), s1 = mk<R>(0, 0);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        s0 = s0 + qa[m] * v[m];
        s1 = s1 + qb[m] * v[m];
      }
      // sums over the TT lanes of this transform
#pragma unroll
      for (int o = 1; o < TT; o <<= 1) {
        s0.re += __shfl_xor(s0.re, o, TT);
        s0.im += __shfl_xor(s0.im, o, TT);
        s1.re += __shfl_xor(s1.re, o, TT);
        s1.im += __shfl_xor(s1.im, o, TT);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = v[m] + fa[m] * s0 + fb[m] * s1;
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
// along y of Tp[a][z][0..ny)[cx] placed at padyl, times the conjugate
// column-stage propagator, summed over z; support rows -> Sp[a][cx][cy].
// The propagator is the plan's table gt[k][cx][cy] when present (with the
// fused z-prefilter terms of fused plans), else e^{i k_z(c,cy) z}.  One
// column per transform slot.
template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_col(const cpx<R> *__restrict__ Tp, cpx<R> *__restrict__ Sp,
              const cpx<R> *__restrict__ twy, const double *__restrict__ prop_a,
              Geom g, int ldt, const cpx<R> *__restrict__ gt) {
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
    const cpx<R> *gk = gt ? gt + ((size_t)k * g.Lx + cs) * g.Ly : nullptr;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if (!cyv[m]) continue;
      cpx<R> e;
      if (gk) {
        const cpx<R> q = gk[cyc[m]];
        e = mk<R>(q.re, -q.im);
      } else {
        e = phasor(-pa[cyc[m]] * z, R(0));
      }
      acc[m] = acc[m] + v[m] * e;
    }
  }
  if (!live) return;
  cpx<R> *dst = Sp + ((size_t)ai * g.Lx + c) * g.Ly;
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if (cyv[m]) dst[cyc[m]] = acc[m];
}

// k_adj_col with the slices' T rows staged through LDS: a block takes CPB
// consecutive support columns (rows of 64 or 128 bytes, read whole) in NP
// passes of TPB transforms; slice k + 1's tile is loaded into registers
// while slice k is transformed.  The sums are the same, in the same order,
// as k_adj_col's.
template <int N>
struct AdjColTile {
  static constexpr int TT = N / 8;
  static constexpr int TPB = 256 / TT;
  static constexpr int CPB0 = TPB > 16 ? TPB : 16;
  static constexpr int CPB = CPB0 < 2 * TPB ? CPB0 : 2 * TPB;  // columns per block
  static constexpr int NP = CPB / TPB;                          // passes
  static constexpr int NE = 12;  // tile elements per thread at most
};

template <typename R, int N>
__global__ void __launch_bounds__(256)
    k_adj_col_t(const cpx<R> *__restrict__ Tp, cpx<R> *__restrict__ Sp,
                const cpx<R> *__restrict__ twy, const double *__restrict__ prop_a,
                Geom g, int ldt, const cpx<R> *__restrict__ gt) {
  using FFT = Fft8<R, N, -1>;
  using CT = AdjColTile<N>;
  constexpr int TT = CT::TT, TPB = CT::TPB, CPB = CT::CPB, NP = CT::NP, NE = CT::NE;
  static_assert(TT <= 64, "one transform per wavefront at most");
  const int t = threadIdx.x / TT, j = threadIdx.x - t * TT;
  cpx<R> *lds = reinterpret_cast<cpx<R> *>(odt_smem);
  cpx<R> *b0 = lds + t * FFT::PADN;
  cpx<R> *tile = lds + TPB * FFT::PADN;  // ny rows x (CPB + 1)
  const int c0 = blockIdx.x * CPB, ai = blockIdx.y;
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
  const int ne = ny * CPB;  // <= NE * 256 (the launcher checks)
  // tile element e = threadIdx.x + 256 u: row e / CPB, column e % CPB
  cpx<R> pf[NE];
  auto fetch = [&](int k) {
    const cpx<R> *src = Tp + (((size_t)ai * ln + k) * ny) * ldt + c0;
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = (int)threadIdx.x + 256 * u;
      const int r = e / CPB, cl = e - r * CPB;
      pf[u] = e < ne && c0 + cl < g.Lx ? src[(size_t)r * ldt + cl] : mk<R>(0, 0);
    }
  };
  cpx<R> acc[NP][8];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[q][m] = mk<R>(0, 0);
  fetch(0);
  for (int k = 0; k < ln; ++k) {
    __syncthreads();  // the previous slice's tile reads are done
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = (int)threadIdx.x + 256 * u;
      if (e < ne) tile[(e / CPB) * (CPB + 1) + e % CPB] = pf[u];
    }
    __syncthreads();
    if (k + 1 < ln) fetch(k + 1);
    const double z = g.z0 + (double)k;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int cl = q * TPB + t, c = c0 + cl;
      const int cs = c < g.Lx ? c : 0;
      cpx<R> v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int y = j + m * TT - g.padyl;
        v[m] = (y >= 0 && y < ny) ? tile[y * (CPB + 1) + cl] : mk<R>(0, 0);
      }
      FFT::run(v, b0, nullptr, tw, j);
      const cpx<R> *gk = gt ? gt + ((size_t)k * g.Lx + cs) * g.Ly : nullptr;
      const double *pa = prop_a + (size_t)cs * g.Ly;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (!cyv[m]) continue;
        cpx<R> e;
        if (gk) {
          const cpx<R> qv = gk[cyc[m]];
          e = mk<R>(qv.re, -qv.im);
        } else {
          e = phasor(-pa[cyc[m]] * z, R(0));
        }
        acc[q][m] = acc[q][m] + v[m] * e;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int c = c0 + q * TPB + t;
    if (c >= g.Lx) continue;
    cpx<R> *dst = Sp + ((size_t)ai * g.Lx + c) * g.Ly;
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (cyv[m]) dst[cyc[m]] = acc[q][m];
  }
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
// The same transposed gather from a block table of forward taps.  The
// gradient comes transposed to planes-innermost, gt[z][x][y], so the four
// lanes that share a coefficient (each taking YC/4 consecutive planes) read
// one 128-byte line per candidate voxel: a wave's load touches one line per
// coefficient instead of one per lane and plane.  Per block (a 16 x 16
// coefficient tile, YC planes, one angle) the voxels of the tile's
// candidate box get their forward tap data once -- tap start (packed) and
// the per-axis weights, formed exactly as the forward kernel forms them --
// and the candidate search reads them from LDS.
// ---------------------------------------------------------------------------
template <typename R, int ORDER>
struct AdjBox {
  static constexpr int K = ORDER / 2 + 1;  // taps reach at most K past an edge
  // box side: the candidate ranges of a tile's (virtual) coefficient
  // rectangle, at most 15 + 2K wide, rotated (factor <= sqrt 2), plus the
  // candidate square (2h = ORDER + 1, also rotated) and the margins
  static constexpr int S = (int)(1.41422 * (15 + 2 * K + ORDER + 1)) + 5;
  static constexpr int N = S * S;
  static constexpr int INVALID = 0x7fff7fff;  // start 32767: no coefficient matches
  static constexpr size_t LDS = (size_t)N * (sizeof(int) + 2 * (ORDER + 1) * sizeof(R));
  // block tables up to 100 KiB (float64 orders 4 and 5 keep k_adj_rotate)
  static constexpr bool FITS = LDS <= 100 * 1024;
};

// Forward tap data of output voxel (z, x) (k_rotate_lds2's prep): packed
// start (z in the low, x in the high 16 bits; INVALID outside the plane) and
// the weights of taps 0 .. ORDER along each axis.
template <typename R, int ORDER>
__device__ __forceinline__ int adj_voxel_taps(int z, int x, double m00, double m01,
                                              double m10, double m11, double oz,
                                              double ox, int nz, int nx, R *wz, R *wx) {
  const double cz = affine2((double)z, (double)x, m00, m01, oz);
  const double cx = affine2((double)z, (double)x, m10, m11, ox);
  const bool valid = cz >= 0.0 && cz <= (double)(nz - 1) && cx >= 0.0 &&
                     cx <= (double)(nx - 1);
  const double fz = (ORDER & 1) ? floor(cz) : floor(cz + 0.5);
  const double fx = (ORDER & 1) ? floor(cx) : floor(cx + 0.5);
  spline_weights_t<ORDER, R>((R)(cz - fz), wz);
  spline_weights_t<ORDER, R>((R)(cx - fx), wx);
  const int sz = (int)fz - ORDER / 2, sx = (int)fx - ORDER / 2;
  return valid ? ((sz & 0xffff) | (sx << 16)) : AdjBox<R, ORDER>::INVALID;
}

template <typename R, int ORDER, int YC>
__global__ void __launch_bounds__(256)
    k_adj_rotate_t(const cpx<R> *__restrict__ gt, const double *__restrict__ params,
                   cpx<R> *__restrict__ slab, int ny, int nz, int nx) {
  using Box = AdjBox<R, ORDER>;
  constexpr int NW = ORDER + 1, K = Box::K;
  constexpr int PPL = YC / 4;  // planes per lane, 32 bytes
  static_assert(PPL * sizeof(cpx<R>) == 32, "four lanes per 128-byte line");
  __shared__ int bs[Box::N];
  __shared__ R bwz[NW][Box::N];
  __shared__ R bwx[NW][Box::N];
  const int ntx = (nx + 15) >> 4;
  const int tz = blockIdx.x / ntx, tx = blockIdx.x - tz * ntx;
  const int a = blockIdx.y, y0 = blockIdx.z * YC;
  const double *p = params + 6 * (size_t)a;
  const double m00 = p[0], m01 = p[1], m10 = p[2], m11 = p[3];
  const double oz = p[4], ox = p[5];
  constexpr double h = 0.5 * (ORDER + 1);
  const double hz = h * (fabs(m00) + fabs(m10)) + 1e-7;
  const double hx = h * (fabs(m01) + fabs(m11)) + 1e-7;
  const int nzc = (int)floor(2.0 * hz) + 1, nxc = (int)floor(2.0 * hx) + 1;
  // the tile's coefficient rectangle, mirrored virtual positions included
  const int tz0 = tz * 16, tz1 = min(tz0 + 15, nz - 1);
  const int tx0 = tx * 16, tx1 = min(tx0 + 15, nx - 1);
  const int Z0 = tz0 <= K ? -K : tz0, Z1 = tz1 >= nz - 1 - K ? nz - 1 + K : tz1;
  const int X0 = tx0 <= K ? -K : tx0, X1 = tx1 >= nx - 1 - K ? nx - 1 + K : tx1;
  // candidate box: the candidate ranges are monotone in the coefficient
  // position, so the rectangle's corners bound them (one sample of margin)
  int bz0 = 0, bz1 = 0, bx0 = 0, bx1 = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double dz = (double)((c & 1) ? Z1 : Z0) - oz;
    const double dx = (double)((c & 2) ? X1 : X0) - ox;
    const int zl = (int)ceil(m00 * dz + m10 * dx - hz);
    const int xl = (int)ceil(m01 * dz + m11 * dx - hx);
    bz0 = c ? min(bz0, zl) : zl;
    bz1 = c ? max(bz1, zl + nzc - 1) : zl + nzc - 1;
    bx0 = c ? min(bx0, xl) : xl;
    bx1 = c ? max(bx1, xl + nxc - 1) : xl + nxc - 1;
  }
  bz0 = max(bz0 - 1, 0);
  bz1 = min(bz1 + 1, nz - 1);
  bx0 = max(bx0 - 1, 0);
  bx1 = min(bx1 + 1, nx - 1);
  const int bh = bz1 - bz0 + 1, bw = bx1 - bx0 + 1;
  const int nbox = bh > 0 && bw > 0 ? bh * bw : 0;
  const bool table = nbox <= Box::N;  // else every candidate computes its taps
  if (table)
    for (int e = threadIdx.x; e < nbox; e += 256) {
      const int zz = bz0 + e / bw, xx = bx0 + e % bw;
      R wz[NW], wx[NW];
      bs[e] = adj_voxel_taps<R, ORDER>(zz, xx, m00, m01, m10, m11, oz, ox, nz, nx, wz, wx);
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        bwz[j][e] = wz[j];
        bwx[j][e] = wx[j];
      }
    }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int yb = y0 + (lane & 3) * PPL;  // this lane's first plane
  constexpr int NONE = -(1 << 30);
  for (int it = 0; it < 4; ++it) {
    const int jz = tz0 + it * 4 + wave, jx = tx0 + (lane >> 2);
    if (jz > tz1 || jx > tx1) continue;
    const int vz1 = (jz >= 1 && jz <= K)                    ? -jz
                    : (jz <= nz - 2 && jz >= nz - 1 - K) ? 2 * nz - 2 - jz
                                                          : NONE;
    const int vx1 = (jx >= 1 && jx <= K)                    ? -jx
                    : (jx <= nx - 2 && jx >= nx - 1 - K) ? 2 * nx - 2 - jx
                                                          : NONE;
    cpx<R> acc[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) acc[k] = mk<R>(0, 0);
    const bool lane_on = yb < ny;  // planes beyond ny: nothing to read
    for (int iv = 0; iv < 4 && nbox > 0; ++iv) {
      const int vz = (iv & 1) ? vz1 : jz, vx = (iv & 2) ? vx1 : jx;
      if (vz == NONE || vx == NONE) continue;
      const double dz = (double)vz - oz, dx = (double)vx - ox;
      const int zlo = (int)ceil(m00 * dz + m10 * dx - hz);
      const int xlo = (int)ceil(m01 * dz + m11 * dx - hx);
      // this lane's PPL planes of voxel (z, x): 32 bytes (ny is a multiple
      // of 4)
      auto add = [&](int z, int x, R w) {
        const float4 *src =
            reinterpret_cast<const float4 *>(gt + ((size_t)z * nx + x) * ny + yb);
        const float4 q0 = src[0], q1 = src[1];
        cpx<R> v[PPL];
        static_assert(sizeof(v) == 2 * sizeof(float4), "32-byte lane block");
        __builtin_memcpy(&v[0], &q0, sizeof(float4));
        __builtin_memcpy(reinterpret_cast<char *>(v) + sizeof(float4), &q1, sizeof(float4));
#pragma unroll
        for (int k = 0; k < PPL; ++k) acc[k] = acc[k] + w * v[k];
      };
      if (table) {
        // candidates in the volume lie in the box (the box is clipped to
        // the volume): one table read per candidate, one branch for the
        // accepted ones
        for (int iz = 0; iz < nzc; ++iz) {
          const int ez = zlo + iz - bz0;
          for (int ix = 0; ix < nxc; ++ix) {
            const int ex = xlo + ix - bx0;
            const bool inb = (unsigned)ez < (unsigned)bh && (unsigned)ex < (unsigned)bw;
            const int e = inb ? ez * bw + ex : 0;
            const int pk = bs[e];
            const int tzi = vz - (int)(short)(pk & 0xffff), txi = vx - (pk >> 16);
            const bool ok = inb && lane_on && (unsigned)tzi <= (unsigned)ORDER &&
                            (unsigned)txi <= (unsigned)ORDER;
            if (ok) add(bz0 + ez, bx0 + ex, bwz[tzi][e] * bwx[txi][e]);
          }
        }
      } else {
        for (int iz = 0; iz < nzc; ++iz) {
          const int z = zlo + iz;
          if (z < 0 || z >= nz) continue;
          for (int ix = 0; ix < nxc; ++ix) {
            const int x = xlo + ix;
            if (x < 0 || x >= nx) continue;
            R wz[NW], wx[NW];
            const int pk = adj_voxel_taps<R, ORDER>(z, x, m00, m01, m10, m11, oz, ox, nz,
                                                    nx, wz, wx);
            const int tzi = vz - (int)(short)(pk & 0xffff), txi = vx - (pk >> 16);
            if (!lane_on || (unsigned)tzi > (unsigned)ORDER || (unsigned)txi > (unsigned)ORDER)
              continue;
            R wzv = R(0), wxv = R(0);
#pragma unroll
            for (int j = 0; j < NW; ++j) {
              wzv = j == tzi ? wz[j] : wzv;
              wxv = j == txi ? wx[j] : wxv;
            }
            add(z, x, wzv * wxv);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k)
      if (yb + k < ny)
        slab[(((size_t)a * ny + yb + k) * nz + jz) * nx + jx] = acc[k];
  }
}

// g[z][y][x] -> gt[z][x][y], 32 x 32 tiles through LDS.
template <typename R>
__global__ void __launch_bounds__(256)
    k_adj_transpose(const cpx<R> *__restrict__ g, cpx<R> *__restrict__ gt, int ny, int nx) {
  __shared__ cpx<R> t[32][33];
  const int z = blockIdx.z, by = blockIdx.y * 32, bx = blockIdx.x * 32;
  const int c = threadIdx.x & 31, r0 = threadIdx.x >> 5;
  for (int r = r0; r < 32; r += 8) {
    const int y = by + r, x = bx + c;
    if (y < ny && x < nx) t[r][c] = g[((size_t)z * ny + y) * nx + x];
  }
  __syncthreads();
  for (int r = r0; r < 32; r += 8) {
    const int x = bx + r, y = by + c;
    if (y < ny && x < nx) gt[((size_t)z * nx + x) * ny + y] = t[c][r];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R, int ORDER>
static hipError_t adj_rotate_launch(const cpx<R> *g, int nb, const double *params,
                                    cpx<R> *slab, int ny, int nz, int nx,
                                    hipStream_t st) {
  constexpr int YC = sizeof(R) == 4 ? 16 : 8;
  dim3 grid(((nz + 15) / 16) * ((nx + 15) / 16), nb, (ny + YC - 1) / YC);
  k_adj_rotate<R, ORDER, YC><<<grid, 256, 0, st>>>(g, params, slab, ny, nz, nx);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_adj_rotate(int order, const cpx<R> *g, int nb,
                             const double *params, cpx<R> *slab, int ny, int nz,
                             int nx, hipStream_t st) {
  if (nz < 8 || nx < 8) return hipErrorInvalidValue;
  switch (order) {
    case 2: return adj_rotate_launch<R, 2>(g, nb, params, slab, ny, nz, nx, st);
    case 3: return adj_rotate_launch<R, 3>(g, nb, params, slab, ny, nz, nx, st);
    case 4: return adj_rotate_launch<R, 4>(g, nb, params, slab, ny, nz, nx, st);
    case 5: return adj_rotate_launch<R, 5>(g, nb, params, slab, ny, nz, nx, st);
  }
  return hipErrorInvalidValue;
}

template <typename R, int ORDER>
static hipError_t adj_rotate_t_launch(const cpx<R> *gt, int nb, const double *params,
                                      cpx<R> *slab, int ny, int nz, int nx,
                                      hipStream_t st) {
  if constexpr (AdjBox<R, ORDER>::FITS) {
    constexpr int YC = 32 / (int)sizeof(cpx<R>) * 4;
    dim3 grid(((nz + 15) / 16) * ((nx + 15) / 16), nb, (ny + YC - 1) / YC);
    k_adj_rotate_t<R, ORDER, YC><<<grid, 256, 0, st>>>(gt, params, slab, ny, nz, nx);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <typename R>
bool adj_rotate_t_supported(int order, int ny, int nz, int nx) {
  const bool fits = order == 2   ? AdjBox<R, 2>::FITS
                    : order == 3 ? AdjBox<R, 3>::FITS
                    : order == 4 ? AdjBox<R, 4>::FITS
                    : order == 5 ? AdjBox<R, 5>::FITS
                                 : false;
  return fits && nz >= 8 && nx >= 8 && ny % 4 == 0 && nz <= 32000 && nx <= 32000 &&
         nz <= 65535;
}

template <typename R>
hipError_t launch_adj_rotate_t(int order, const cpx<R> *gt, int nb, const double *params,
                               cpx<R> *slab, int ny, int nz, int nx, hipStream_t st) {
  if (!adj_rotate_t_supported<R>(order, ny, nz, nx)) return hipErrorInvalidValue;
  switch (order) {
    case 2: return adj_rotate_t_launch<R, 2>(gt, nb, params, slab, ny, nz, nx, st);
    case 3: return adj_rotate_t_launch<R, 3>(gt, nb, params, slab, ny, nz, nx, st);
    case 4: return adj_rotate_t_launch<R, 4>(gt, nb, params, slab, ny, nz, nx, st);
    case 5: return adj_rotate_t_launch<R, 5>(gt, nb, params, slab, ny, nz, nx, st);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_adj_transpose(const cpx<R> *g, cpx<R> *gt, int ny, int nz, int nx,
                                hipStream_t st) {
  if (nz > 65535) return hipErrorInvalidValue;
  dim3 grid((nx + 31) / 32, (ny + 31) / 32, nz);
  k_adj_transpose<R><<<grid, 256, 0, st>>>(g, gt, ny, nx);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_adj_fir(const cpx<R> *in, cpx<R> *out, long long nlines, int n,
                          long long n0, long long s0, long long s1, long long es,
                          const R *hz, int K, hipStream_t st) {
  if (nlines <= 0 || n < 1) return hipSuccess;
  if (n0 < 1) return hipErrorInvalidValue;
  const long long ngrp = (nlines + 255) / 256;
  const int nzb = (n + kFirOB - 1) / kFirOB;
  const long long blocks = ((ngrp + 7) / 8) * 8 * nzb;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  k_adj_fir<R><<<(unsigned)blocks, 256, 0, st>>>(in, out, nlines, n, n0, s0, s1, es, hz,
                                                 K, nzb);
  return hipGetLastError();
}

int adj_fir_pad() { return adj_fir_pad_c; }

template <typename R>
hipError_t launch_adj_fir_rows(const cpx<R> *in, cpx<R> *out, long long nrows, int n,
                               const R *hz, int K, hipStream_t st) {
  if (nrows <= 0 || n < 1) return hipSuccess;
  const size_t lds = 4 * (size_t)(n + 2 * K) * sizeof(cpx<R>);
  if (lds > 65536) return hipErrorInvalidValue;
  k_adj_fir_rows<R><<<(unsigned)((nrows + 3) / 4), 256, lds, st>>>(in, out, nrows, n,
                                                                  hz, K);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_adj_prefilter_rows(cpx<R> *rows, long long nrows, int n,
                                     int order, hipStream_t st) {
  int np = 0;
  double z1 = 0, z2 = 0, gain = 1;
  if (!iir_poles(order, &np, &z1, &z2, &gain)) return hipErrorInvalidValue;
  if (nrows <= 0) return hipSuccess;
  const size_t row_bytes = (size_t)(2 * n + 2) * sizeof(R);
  const int rb = (int)std::max<size_t>(1, std::min<size_t>(32, 65536 / row_bytes));
  const size_t lds = rb * row_bytes;
  if (lds > 65536) return hipErrorInvalidValue;
  const long long blocks = (nrows + rb - 1) / rb;
  k_adj_iir_rows<R><<<(unsigned)blocks, 64, lds, st>>>(
      (R *)rows, nrows, n, rb, np, z1, z2, gain);
  return hipGetLastError();
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
                                           (const cpx<R> *)tw, g, rb, ldt,
                                           (const R *)aux);
  } else if (stage == 1) {
    using CT = AdjColTile<N>;
    const size_t sht =
        ((size_t)CT::TPB * lds_padded(N) + (size_t)g.Ny * (CT::CPB + 1)) * sizeof(cpx<R>);
    if ((size_t)g.Ny * CT::CPB <= (size_t)CT::NE * 256 && sht <= 160 * 1024 &&
        !std::getenv("ODTBRAIN_AMD_ADJ_COL_DIRECT")) {
      if (sht > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_adj_col_t<R, N>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)sht);
        if (e != hipSuccess) return e;
      }
      dim3 grid((g.Lx + CT::CPB - 1) / CT::CPB, na);
      k_adj_col_t<R, N><<<grid, 256, sht, st>>>((const cpx<R> *)in, (cpx<R> *)out,
                                                (const cpx<R> *)tw, prop_a, g, ldt,
                                                (const cpx<R> *)aux);
      return hipGetLastError();
    }
    dim3 grid((g.Lx + tpb - 1) / tpb, na);
    k_adj_col<R, N><<<grid, 256, sh, st>>>((const cpx<R> *)in, (cpx<R> *)out,
                                           (const cpx<R> *)tw, prop_a, g, ldt,
                                           (const cpx<R> *)aux);
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

// stage 0: k_adj_row (n = lNx, aux = q0x for fused plans), 1: k_adj_col
// (n = lNy, aux = the propagator table or null), 2: k_adj_inv_col (n = lNy,
// aux = P), 3: k_adj_inv_row (n = lNx).
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
  template hipError_t launch_adj_fir_rows<R>(const cpx<R> *, cpx<R> *,         \
                                             long long, int, const R *, int,    \
                                             hipStream_t);                      \
  template hipError_t launch_adj_fir<R>(const cpx<R> *, cpx<R> *, long long,   \
                                        int, long long, long long, long long,  \
                                        long long, const R *, int,             \
                                        hipStream_t);                          \
  template hipError_t launch_adj_prefilter_rows<R>(cpx<R> *, long long, int,  \
                                                   int, hipStream_t);          \
  template hipError_t launch_adj_rotate<R>(int, const cpx<R> *, int,            \
                                           const double *, cpx<R> *, int, int,  \
                                           int, hipStream_t);                   \
  template bool adj_rotate_t_supported<R>(int, int, int, int);                  \
  template hipError_t launch_adj_rotate_t<R>(int, const cpx<R> *, int,          \
                                             const double *, cpx<R> *, int, int, \
                                             int, hipStream_t);                  \
  template hipError_t launch_adj_transpose<R>(const cpx<R> *, cpx<R> *, int, int, \
                                              int, hipStream_t);                  \
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

// Sinogram preparation (odtbrain/_prepare_sino.py:7-182): the Rytov (or
// plain phase) sinogram of a complex field, with phase unwrapping and
// alignment.
//
//   k_prep_phase       angle(u / u0) and log|u / u0| per sample (numpy's
//                      Smith complex division when u0 is an array)
//   k_unwrap_rows      2D sinograms: numpy.unwrap along the detector axis,
//                      one thread per row, the reference's sequential sum
//   3D sinograms: every angle's image unwrapped by the reliability-sorted
//   algorithm of Herraez et al. (Applied Optics 41 (2002) 7437), the
//   algorithm of skimage.restoration.unwrap_phase (_prepare_sino.py:172):
//   k_uw_reliability   wrapped second differences per pixel (borders: 1e7)
//   k_uw_edges         4-neighbour edges, key = sum of the two reliabilities
//   segmented radix sort (rocPRIM, stable: equal keys keep creation order)
//   k_uw_merge_lds     images of at most 65536 pixels: one wavefront per
//                      image merges groups edge by edge in sorted order
//                      (union by size with whole-turn offsets, the
//                      published merge rules), group sizes in LDS, the
//                      root finds of 64 edges at a time across the lanes
//   k_uw_gather        larger images: one thread per image, the groups in
//                      global memory; a helper wavefront warms L2 with the
//                      nodes of the edges just ahead
//   k_uw_apply         value + 2 pi * increment
//   k_min_*, k_align   align_unwrapped: steps of the unwrapped corner
//                      samples along the angles, their mode per angle,
//                      whole turns from the minimum (divmod_neg)
//   k_finish           u0 * (log|.| + i (phase - remove)) or phase - remove
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "odt_device.h"
#include "odt_launch.h"

namespace odt {

namespace {
constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kBorderReliability = 1e7;
}  // namespace

// numpy's complex division (Smith's method, as in its loops), no FMA
template <typename R>
__device__ __forceinline__ cpx<R> np_cdiv(cpx<R> a, cpx<R> b) {
#pragma clang fp contract(off)
  const R ar = fabs(b.re), ai = fabs(b.im);
  if (ar >= ai) {
    if (ar == R(0) && ai == R(0)) return mk<R>(a.re / ar, a.im / ai);
    const R rat = b.im / b.re, scl = R(1) / (b.re + b.im * rat);
    return mk<R>((a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl);
  }
  const R rat = b.re / b.im, scl = R(1) / (b.im + b.re * rat);
  return mk<R>((a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl);
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_prep_phase(const cpx<R> *__restrict__ u, const cpx<R> *__restrict__ u0,
                 long long n, R *__restrict__ phase, R *__restrict__ lna) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cpx<R> q = u[i];
  if (u0) q = np_cdiv(q, u0[i]);
  phase[i] = atan2(q.im, q.re);
  if (lna) lna[i] = log(hypot(q.re, q.im));
}

// numpy.mod for floats (result takes the divisor's sign)
template <typename R>
__device__ __forceinline__ R np_mod(R a, R b) {
  R m = fmod(a, b);
  if (m != R(0)) {
    if ((b < R(0)) != (m < R(0))) m += b;
  } else {
    m = copysign(R(0), b);
  }
  return m;
}

// numpy.unwrap (period 2 pi, discont pi) of n values at stride s, in place:
// corrections from the differences of the original values, summed in order.
template <typename R>
__device__ void np_unwrap_line(R *p, int n, long long s) {
#pragma clang fp contract(off)
  const R pi = (R)kPi, period = (R)kTwoPi;
  R prev = p[0], cum = R(0);
  for (int i = 1; i < n; ++i) {
    const R cur = p[(long long)i * s];
    const R dd = cur - prev;
    R ddmod = np_mod<R>(dd - (-pi), period) + (-pi);
    if (ddmod == -pi && dd > R(0)) ddmod = pi;
    R pc = ddmod - dd;
    if (fabs(dd) < pi) pc = R(0);
    cum = cum + pc;
    p[(long long)i * s] = cur + cum;
    prev = cur;
  }
}

template <typename R>
__global__ void k_unwrap_rows(R *phase, int rows, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) np_unwrap_line<R>(phase + (long long)r * n, n, 1);
}

__device__ __forceinline__ double uw_wrap(double d) {
  if (d > kPi) return d - kTwoPi;
  if (d < -kPi) return d + kTwoPi;
  return d;
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_uw_reliability(const R *__restrict__ phase, int A, int H, int W,
                     double *__restrict__ rel) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npix = (long long)H * W;
  if (i >= (long long)A * npix) return;
  const long long a = i / npix;
  const int p = (int)(i - a * npix), y = p / W, x = p - y * W;
  if (y == 0 || y == H - 1 || x == 0 || x == W - 1) {
    rel[i] = kBorderReliability;
    return;
  }
  const R *im = phase + a * npix;
  auto v = [&](int yy, int xx) { return (double)im[yy * W + xx]; };
  const double c = v(y, x);
  const double h = uw_wrap(v(y, x - 1) - c) - uw_wrap(c - v(y, x + 1));
  const double vv = uw_wrap(v(y - 1, x) - c) - uw_wrap(c - v(y + 1, x));
  const double d1 = uw_wrap(v(y - 1, x - 1) - c) - uw_wrap(c - v(y + 1, x + 1));
  const double d2 = uw_wrap(v(y - 1, x + 1) - c) - uw_wrap(c - v(y + 1, x - 1));
  rel[i] = h * h + vv * vv + d1 * d1 + d2 * d2;
}

__device__ __forceinline__ int uw_find_wrap(double v1, double v2) {
  const double d = v1 - v2;
  if (d > kPi) return -1;
  if (d < -kPi) return 1;
  return 0;
}

// Edge e of an image: e < H (W - 1) horizontal (y, x)-(y, x + 1) row by
// row, then vertical (y, x)-(y + 1, x).
__device__ __forceinline__ void uw_edge(int e, int H, int W, int *p1, int *p2) {
  const int nh = H * (W - 1);
  if (e < nh) {
    const int y = e / (W - 1), x = e - y * (W - 1);
    *p1 = y * W + x;
    *p2 = *p1 + 1;
  } else {
    *p1 = e - nh;
    *p2 = *p1 + W;
  }
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_uw_edges(const double *__restrict__ rel, const R *__restrict__ phase, int A,
               int H, int W, double *__restrict__ keys, int *__restrict__ vals,
               unsigned *__restrict__ offsets) {
  const long long E = (long long)H * (W - 1) + (long long)(H - 1) * W;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= A && i >= 0) {
    if (i <= A) offsets[i] = (unsigned)(i * E);
  }
  if (i >= (long long)A * E) return;
  const long long a = i / E;
  const int e = (int)(i - a * E);
  int p1, p2;
  uw_edge(e, H, W, &p1, &p2);
  const double *r = rel + a * (long long)H * W;
  keys[i] = r[p1] + r[p2];
  // edge id and its whole-turn step (find_wrap + 1 in the low two bits)
  const R *img = phase + a * (long long)H * W;
  vals[i] = (e << 2) | (uw_find_wrap((double)img[p1], (double)img[p2]) + 1);
}

// Union-find node per pixel: a root holds (-size of its group, its own
// increment), any other node (parent, its increment relative to the
// parent).  find: root of p and p's increment, compressing the path.
__device__ __forceinline__ int uw_find(int p, int2 *node, int *inc, int2 *root) {
  int r = p, s = 0;
  int2 nr = node[r];
  while (nr.x >= 0) {
    s += nr.y;
    r = nr.x;
    nr = node[r];
  }
  // compress: every node on the path points at r with its total offset
  int q = p, sq = s;
  while (q != r) {
    const int2 nq = node[q];
    if (nq.x == r) break;
    node[q] = make_int2(r, sq);
    sq -= nq.y;
    q = nq.x;
  }
  *inc = s + nr.y;
  *root = nr;
  return r;
}

// One image per block: one thread merges the groups of the published
// algorithm as a union-find forest (union by size), edge by edge in sorted
// order; the sorted edges carry their whole-turn step and are read 16 at a
// time ahead of use.  A second wavefront walks at most kUwAhead edges
// ahead of the merging thread and touches the nodes of those edges with
// agent-scope loads: they fill L2 (the node arrays of all images exceed
// it) but not the merging wavefront's L1, so no stale line can reach the
// merges; the merging thread never waits for it.  The helper's wait ends
// because the merging thread always runs to the last edge and then
// publishes that.
constexpr int kUwAhead = 32;
constexpr int kUwThreads = 128;

template <typename R>
__global__ void __launch_bounds__(kUwThreads)
    k_uw_gather(int A, int H, int W, const int *__restrict__ sorted,
                int2 *node_all) {
  const int a = blockIdx.x;
  if (a >= A) return;  // uniform per block
  const long long npix = (long long)H * W;
  const long long E = (long long)H * (W - 1) + (long long)(H - 1) * W;
  int2 *node = node_all + a * npix;
  for (long long p = threadIdx.x; p < npix; p += kUwThreads)
    node[p] = make_int2(-1, 0);
  const int *es = sorted + a * E;
  __shared__ long long progress;  // edges the merging thread has finished
  if (threadIdx.x == 0) progress = 0;
  __syncthreads();
  if (threadIdx.x == 64) {
    // batches of HB edges: all 2 HB loads are in flight before any result
    // is used, so the helper outruns the merges it serves
    constexpr int HB = 8;
    int touch = 0;
    for (long long k0 = 0; k0 < E; k0 += HB) {
      while (k0 > __hip_atomic_load(&progress, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP) +
                       kUwAhead)
        __builtin_amdgcn_s_sleep(2);
      int t[2 * HB];
#pragma unroll
      for (int u = 0; u < HB; ++u) {
        const long long k = k0 + u < E ? k0 + u : E - 1;
        int p1, p2;
        uw_edge(es[k] >> 2, H, W, &p1, &p2);
        t[2 * u] = __hip_atomic_load(&node[p1].x, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        t[2 * u + 1] = __hip_atomic_load(&node[p2].x, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int u = 0; u < 2 * HB; ++u) touch += t[u];
    }
    __asm__ volatile("" : : "v"(touch));  // keeps the touching loads
    return;
  }
  if (threadIdx.x != 0) return;
  constexpr int CH = 16;
  for (long long k0 = 0; k0 < E; k0 += CH) {
    __hip_atomic_store(&progress, k0, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
    int ev[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) ev[u] = k0 + u < E ? es[k0 + u] : -1;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (ev[u] < 0) break;
      int p1, p2;
      uw_edge(ev[u] >> 2, H, W, &p1, &p2);
      const int e = (ev[u] & 3) - 1;
      int i1, i2;
      int2 n1, n2;
      const int r1 = uw_find(p1, node, &i1, &n1);
      const int r2 = uw_find(p2, node, &i2, &n2);
      if (r1 == r2) continue;
      const int s1 = -n1.x, s2 = -n2.x;
      if (s2 == 1) {  // p2 alone joins p1's group
        node[p2] = make_int2(r1, i1 - e - n1.y);
        node[r1] = make_int2(-(s1 + 1), n1.y);
      } else if (s1 == 1) {  // p1 alone joins p2's group
        node[p1] = make_int2(r2, i2 + e - n2.y);
        node[r2] = make_int2(-(s2 + 1), n2.y);
      } else if (s1 > s2) {  // group 2 shifted onto group 1
        node[r2] = make_int2(r1, n2.y + (i1 - e - i2) - n1.y);
        node[r1] = make_int2(-(s1 + s2), n1.y);
      } else {  // group 1 shifted onto group 2
        node[r1] = make_int2(r2, n1.y + (i2 + e - i1) - n2.y);
        node[r2] = make_int2(-(s1 + s2), n2.y);
      }
    }
  }
  __hip_atomic_store(&progress, E, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Images of at most 65536 pixels: the same merges in the same order, with
// the group sizes in LDS.  Three facts of the merge rules above carry it:
// a root's own increment stays 0 (every node starts at 0 and a kept root
// keeps its value), so a merge needs only each endpoint's increment
// relative to its root; the increment of a pixel relative to any of its
// ancestors never changes afterwards (a merge shifts a whole group), so a
// link read at any time stays valid; and union by size keeps every pixel
// within 16 links of its root.
//   gsz[p] (LDS, 16 bits): p's group size less one while p is a root,
//   kUwNotRoot once it is not; kUwSat once the group holds more than half
//   the pixels (it then wins every size comparison, as its true size would).
// Per batch of 64 sorted edges, lane k of one wave walks both ends of edge k
// to their roots through the global links (pointing each node it passes at
// its grandparent: only nodes with a link are written, never by a merge),
// from the first links loaded a batch earlier; the other wave then merges
// the batch, in rounds of conflict-free edges and the rest in edge order,
// while the next batch is walked.  A merge writes the demoted root's link
// (global) and the two size words (LDS).
constexpr int kUwLdsPix = 65536;
constexpr unsigned short kUwSat = 0xFFFE;  // majority group (more than half the pixels)
constexpr int kUwClaims = 2048;            // claim slots (LDS, after the size words)
constexpr int kUwMinWinners = 2;
constexpr unsigned short kUwNotRoot = 0xFFFF;

extern __shared__ __align__(16) unsigned char uw_smem[];
__host__ __device__ constexpr size_t uw_claims_offset(long long npix) {
  return ((size_t)npix * sizeof(unsigned short) + 15) / 16 * 16;
}

// links as single 64-bit accesses (x = parent or -1 at a root, y = offset)
__device__ __forceinline__ int2 uw_ld(int2 *node, int p) {
  const unsigned long long v =
      __hip_atomic_load(reinterpret_cast<unsigned long long *>(node) + p,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
}
__device__ __forceinline__ void uw_st(int2 *node, int p, int parent, int off) {
  const unsigned long long v =
      (unsigned long long)(unsigned)parent | ((unsigned long long)(unsigned)off << 32);
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(node) + p, v,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Roots of p1 and p2 and their increments relative to them, the two walks
// in step (their loads in flight together); pre1/pre2: the links as loaded
// earlier (x < 0 if the pixel was a root then).
__device__ __forceinline__ void uw_walk2(int p1, int2 pre1, int p2, int2 pre2,
                                         int2 *node, const unsigned short *gsz,
                                         int *r1o, int *s1o, int *r2o, int *s2o) {
  bool d1 = gsz[p1] != kUwNotRoot, d2 = gsz[p2] != kUwNotRoot;
  int r1 = p1, r2 = p2, s1 = 0, s2 = 0;
  int2 v1 = pre1, v2 = pre2;
  if (!d1 && pre1.x < 0) v1 = uw_ld(node, p1);
  if (!d2 && pre2.x < 0) v2 = uw_ld(node, p2);
  // a link that still reads as a root (its merge's store not yet in
  // memory): the walk stops there and the merging side walks on
  d1 = d1 || v1.x < 0;
  d2 = d2 || v2.x < 0;
  for (int d = 0; d < 64 && !(d1 && d2); ++d) {  // at most 16 links each
    // v1/v2: the links of the non-root nodes r1/r2
    bool up1 = false, up2 = false;
    int q1 = 0, q2 = 0;
    if (!d1) {
      q1 = v1.x;
      s1 += v1.y;
      up1 = gsz[q1] == kUwNotRoot;
    }
    if (!d2) {
      q2 = v2.x;
      s2 += v2.y;
      up2 = gsz[q2] == kUwNotRoot;
    }
    int2 w1 = v1, w2 = v2;
    if (up1) w1 = uw_ld(node, q1);
    if (up2) w2 = uw_ld(node, q2);
    if (!d1) {
      const bool go = up1 && w1.x >= 0;
      if (go) uw_st(node, r1, w1.x, v1.y + w1.y);  // r1 skips q1
      d1 = !go;
      r1 = q1;
      v1 = w1;
    }
    if (!d2) {
      const bool go = up2 && w2.x >= 0;
      if (go) uw_st(node, r2, w2.x, v2.y + w2.y);
      d2 = !go;
      r2 = q2;
      v2 = w2;
    }
  }
  *r1o = r1;
  *s1o = s1;
  *r2o = r2;
  *s2o = s2;
}

// Walks on from a root that a merge demoted to the current root, adding
// the link offsets to t and reading its size word into g (the merging
// wave's own links are in memory: its stores complete before each round
// and batch ends); stops at a link that still reads as a root.
__device__ __forceinline__ void uw_resolve(int2 *node, const unsigned short *gsz, int &a,
                                           int &t, unsigned &g) {
  for (int d = 0; d < 64 && g == kUwNotRoot; ++d) {
    const int2 v = uw_ld(node, a);
    if (v.x < 0) break;
    t += v.y;
    a = v.x;
    g = gsz[a];
  }
}

// Two wavefronts per image: wave 1 walks the ends of batch i to their roots
// (and loads the first links of batch i + 1) while wave 0 merges batch
// i - 1; the walk results pass through an LDS ring of two slots and both
// waves meet at one barrier per batch.  The walks see the merges of the
// batch before only in part: a root they return may since have been demoted
// (its size word reads kUwNotRoot, and the merging wave walks on from it).
__global__ void __launch_bounds__(128)
    k_uw_merge_lds(int A, int H, int W, const int *__restrict__ sorted,
                   int2 *node_all) {
  const int a = blockIdx.x;
  if (a >= A) return;  // uniform per block
  const int npix = H * W;
  const int E = H * (W - 1) + (H - 1) * W;
  int2 *node = node_all + (long long)a * npix;
  unsigned short *gsz = reinterpret_cast<unsigned short *>(uw_smem);
  unsigned *claim = reinterpret_cast<unsigned *>(uw_smem + uw_claims_offset(npix));
  int *ring = reinterpret_cast<int *>(claim + kUwClaims);  // [2][5][64]
  const int lane = threadIdx.x & 63;
  const bool walker = threadIdx.x >= 64;  // wave 1
  for (int p = threadIdx.x; p < npix; p += 128) {
    uw_st(node, p, -1, 0);
    gsz[p] = 0;
  }
  __syncthreads();
  const int *es = sorted + (long long)a * E;
  auto ends = [&](int ev, int &q1, int &q2) {
    q1 = q2 = 0;
    if (ev >= 0) uw_edge(ev >> 2, H, W, &q1, &q2);
  };
  const int nb = (E + 63) / 64;
  const unsigned thr = (unsigned)npix / 2;
  // the walking wave's state: this batch's and the next one's edges, the
  // ends and their first links (all roots at first)
  int wev = -1, wev1 = -1, p1 = 0, p2 = 0;
  int2 pre1 = make_int2(-1, 0), pre2 = make_int2(-1, 0);
  if (walker) {
    wev = lane < E ? es[lane] : -1;
    wev1 = 64 + lane < E ? es[64 + lane] : -1;
    ends(wev, p1, p2);
  }
  for (int i = 0; i <= nb; ++i) {  // both waves: nb + 1 barriers
    if (walker) {
      if (i < nb) {
        int r1 = 0, r2 = 0, s1 = 0, s2 = 0;
        if (wev >= 0) uw_walk2(p1, pre1, p2, pre2, node, gsz, &r1, &s1, &r2, &s2);
        int *slot = ring + (i & 1) * 320;
        slot[lane] = wev;
        slot[64 + lane] = r1;
        slot[128 + lane] = s1;
        slot[192 + lane] = r2;
        slot[256 + lane] = s2;
        int q1, q2;
        ends(wev1, q1, q2);
        pre1 = wev1 >= 0 ? uw_ld(node, q1) : make_int2(-1, 0);
        pre2 = wev1 >= 0 ? uw_ld(node, q2) : make_int2(-1, 0);
        const int k2 = (i + 2) * 64 + lane;
        wev = wev1;
        wev1 = k2 < E ? es[k2] : -1;
        p1 = q1;
        p2 = q2;
      }
    } else if (i >= 1) {
      const int *slot = ring + ((i - 1) & 1) * 320;
      const int ev = slot[lane];
      int a1 = slot[64 + lane], t1 = slot[128 + lane];
      int a2 = slot[192 + lane], t2 = slot[256 + lane];
      const int e = (ev & 3) - 1;
      // The batch's merges, in rounds: an edge whose ends have different
      // current roots claims each root below the majority size (atomic min
      // of its lane in a hashed table); the edges holding all their claims
      // merge at once.  An earlier edge sharing a root with a later one
      // holds that root's claim, so no merge changes the roots or sizes an
      // earlier pending edge will see; merges into the majority group (more
      // than half the pixels: it wins every size comparison, its size word
      // saturated) commute.  The lowest pending edge always merges; a round
      // with fewer than kUwMinWinners merges hands the rest to the in-order
      // loop below.
      unsigned long long pend = __ballot(ev >= 0 && a1 != a2);
      for (int round = 0; pend != 0 && round < 64; ++round) {
        const bool me = (pend >> lane) & 1ull;
        unsigned g1 = 0, g2 = 0;
        if (me) {
          g1 = gsz[a1];
          g2 = gsz[a2];
          uw_resolve(node, gsz, a1, t1, g1);
          uw_resolve(node, gsz, a2, t2, g2);
        }
        const bool live = me && a1 != a2;
        pend = __ballot(live);
        if (pend == 0) break;
        const int h1 = a1 & (kUwClaims - 1), h2 = a2 & (kUwClaims - 1);
        const bool c1 = live && g1 != kUwSat, c2 = live && g2 != kUwSat;
        if (c1) claim[h1] = ~0u;
        if (c2) claim[h2] = ~0u;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the LDS runs a wave's operations in order
        if (c1) atomicMin(&claim[h1], (unsigned)lane);
        if (c2) atomicMin(&claim[h2], (unsigned)lane);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const bool win = live && (!c1 || claim[h1] == (unsigned)lane) &&
                         (!c2 || claim[h2] == (unsigned)lane);
        const unsigned long long wins = __ballot(win);
        if (__builtin_popcountll(wins) < kUwMinWinners) break;
        if (win) {
          const bool keep1 = g2 == 0 || g1 > g2;
          const int child = keep1 ? a2 : a1, target = keep1 ? a1 : a2;
          const unsigned tot = g1 + g2 + 1;
          gsz[target] = (unsigned short)(tot >= thr ? kUwSat : tot);
          gsz[child] = kUwNotRoot;
          uw_st(node, child, target, keep1 ? t1 - e - t2 : t2 + e - t1);
        }
        pend &= ~wins;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);  // the round's links before the next walks
      }
      // the rest in edge order, lane 0 writing
      while (pend) {
        const int k = __builtin_ctzll(pend);
        pend &= pend - 1;
        int b1 = __builtin_amdgcn_readlane(a1, k), b2 = __builtin_amdgcn_readlane(a2, k);
        int u1 = __builtin_amdgcn_readlane(t1, k), u2 = __builtin_amdgcn_readlane(t2, k);
        const int ek = __builtin_amdgcn_readlane(e, k);
        unsigned g1 = gsz[b1], g2 = gsz[b2];
        if (g1 == kUwNotRoot || g2 == kUwNotRoot) {
          __builtin_amdgcn_s_waitcnt(0);  // this loop's links are in memory
          uw_resolve(node, gsz, b1, u1, g1);
          uw_resolve(node, gsz, b2, u2, g2);
        }
        if (b1 == b2) continue;
        // group 2 joins group 1 when it is a single pixel or the smaller one
        const bool keep1 = g2 == 0 || g1 > g2;
        const int child = keep1 ? b2 : b1, target = keep1 ? b1 : b2;
        const unsigned tot = g1 + g2 + 1;
        if (lane == 0) {
          gsz[target] = (unsigned short)(tot >= thr ? kUwSat : tot);
          gsz[child] = kUwNotRoot;
          uw_st(node, child, target, keep1 ? u1 - ek - u2 : u2 + ek - u1);
        }
      }
    }
    __syncthreads();  // merges and walk results of this step before the next
  }
}

// value + 2 pi * increment, in float64 as the published code, stored as R
template <typename R>
__global__ void __launch_bounds__(256)
    k_uw_apply(R *phase, int A, int H, int W, const int2 *__restrict__ node_all) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npix = (long long)H * W;
  if (i >= (long long)A * npix) return;
  const long long a = i / npix;
  const int2 *node = node_all + a * npix;
  int r = (int)(i - a * npix), s = 0;
  int2 nr = node[r];
  while (nr.x >= 0) {
    s += nr.y;
    r = nr.x;
    nr = node[r];
  }
  s += nr.y;
  phase[i] = (R)((double)phase[i] + kTwoPi * (double)s);
}

// minimum of n values: per-block minima, then one block over them
template <typename R>
__global__ void __launch_bounds__(256)
    k_min_blocks(const R *__restrict__ v, long long n, R *__restrict__ part) {
  __shared__ R sm[256];
  R m = R(INFINITY);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const R x = v[i];
    m = (x < m || isnan(x)) ? x : m;
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const R o = sm[threadIdx.x + s];
      if (o < sm[threadIdx.x] || isnan(o)) sm[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// numpy / Python float divmod quotient, then divmod_neg's move of the
// remainder towards zero (_prepare_sino.py:49-57)
__device__ double divmod_neg_q(double a, double b) {
#pragma clang fp contract(off)
  double mod = fmod(a, b);
  double div = (a - mod) / b;
  if (mod != 0.0) {
    if ((b < 0.0) != (mod < 0.0)) {
      mod += b;
      div -= 1.0;
    }
  } else {
    mod = copysign(0.0, b);
  }
  double q;
  if (div != 0.0) {
    q = floor(div);
    if (div - q > 0.5) q += 1.0;
  } else {
    q = copysign(0.0, a / b);
  }
  const double sr = mod > 0.0 ? 1.0 : (mod < 0.0 ? -1.0 : 0.0);
  if (fabs(mod) > b / 2) q += sr;
  return q;
}

// align_unwrapped: remove[a] = mode over the 5 corner samples of
// (sample - unwrap(sample)) at angle a, plus whole turns from the minimum.
// One block; steps (5 x A) in scratch.
template <typename R>
__global__ void k_align(const R *__restrict__ phase, int A, int ndim, int H,
                        int W, const R *__restrict__ part, int nparts,
                        R *__restrict__ samp, double *__restrict__ steps,
                        double *__restrict__ remove) {
#pragma clang fp contract(off)
  const long long npix = (long long)H * W;
  const int t = threadIdx.x;
  if (t < 5) {
    // 3D: (0,0) (0,-1) (-1,0) (-1,-1) (0,1); 2D: columns 0 1 2 -1 -2
    long long o;
    if (ndim == 3) {
      const int ys[5] = {0, 0, H - 1, H - 1, 0};
      const int xs[5] = {0, W - 1, 0, W - 1, 1};
      o = (long long)ys[t] * W + xs[t];
    } else {
      const int cs[5] = {0, 1, 2, W - 1, W - 2};
      o = cs[t];
    }
    R *s = samp + (long long)t * A;
    for (int a = 0; a < A; ++a) s[a] = phase[a * npix + o];
    np_unwrap_line<R>(s, A, 1);
    for (int a = 0; a < A; ++a)
      steps[(long long)t * A + a] = (double)(R)(phase[a * npix + o] - s[a]);
  }
  __syncthreads();
  __shared__ double qmin;
  if (t == 0) {
    R m = R(INFINITY);
    for (int i = 0; i < nparts; ++i) m = (part[i] < m || isnan(part[i])) ? part[i] : m;
    qmin = divmod_neg_q((double)m, kTwoPi);
  }
  __syncthreads();
  for (int a = t; a < A; a += blockDim.x) {
    // scipy.stats.mode: the most frequent value, the smallest on ties
    double best = 0.0;
    int bestc = 0;
    for (int i = 0; i < 5; ++i) {
      const double v = steps[(long long)i * A + a];
      int c = 0;
      for (int j = 0; j < 5; ++j) c += steps[(long long)j * A + a] == v;
      if (c > bestc || (c == bestc && v < best)) {
        best = v;
        bestc = c;
      }
    }
    remove[a] = best + qmin * kTwoPi;
  }
}

// out = u0 * (lna + i (phase - remove)) (Rytov) or phase - remove (radon)
template <typename R>
__global__ void __launch_bounds__(256)
    k_finish(const R *__restrict__ phase, const R *__restrict__ lna,
             const cpx<R> *__restrict__ u0, const double *__restrict__ remove,
             long long n, long long per_angle, void *out) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R ph = phase[i];
  if (remove) ph = (R)((double)ph - remove[i / per_angle]);
  if (!lna) {
    reinterpret_cast<R *>(out)[i] = ph;
    return;
  }
  cpx<R> r = mk<R>(lna[i], ph);
  if (u0) {
    const cpx<R> w = u0[i];
    r = mk<R>(w.re * r.re - w.im * r.im, w.re * r.im + w.im * r.re);
  }
  reinterpret_cast<cpx<R> *>(out)[i] = r;
}

namespace {
size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
unsigned blocks_of(long long n) { return (unsigned)((n + 255) / 256); }
constexpr int kMinParts = 1024;

struct PrepLayout {
  size_t phase, lna, rel, keys0, keys1, vals0, vals1, offs, parent, off, size,
      part, samp, steps, remove, sort_tmp, total;
  size_t sort_bytes;
};

template <typename R>
hipError_t prep_layout(int ndim, int A, int H, int W, bool rytov, PrepLayout *L) {
  const long long n = (long long)A * H * W;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = al256(o + bytes);
    return at;
  };
  L->phase = take(n * sizeof(R));
  L->lna = rytov ? take(n * sizeof(R)) : 0;
  const long long E = ndim == 3 ? (long long)H * (W - 1) + (long long)(H - 1) * W : 0;
  L->sort_bytes = 0;
  if (ndim == 3) {
    L->rel = take(n * sizeof(double));
    L->keys0 = take(A * E * sizeof(double));
    L->keys1 = take(A * E * sizeof(double));
    L->vals0 = take(A * E * sizeof(int));
    L->vals1 = take(A * E * sizeof(int));
    L->offs = take((A + 1) * sizeof(unsigned));
    L->parent = take(n * sizeof(int2));
    L->off = 0;
    L->size = 0;
    hipError_t e = rocprim::segmented_radix_sort_pairs(
        nullptr, L->sort_bytes, (double *)nullptr, (double *)nullptr, (int *)nullptr,
        (int *)nullptr, (unsigned)(A * E), (unsigned)A, (unsigned *)nullptr,
        (unsigned *)nullptr + 1);
    if (e != hipSuccess) return e;
    L->sort_tmp = take(L->sort_bytes);
  }
  L->part = take(kMinParts * sizeof(R));
  L->samp = take(5 * (size_t)A * sizeof(R));
  L->steps = take(5 * (size_t)A * sizeof(double));
  L->remove = take((size_t)A * sizeof(double));
  L->total = o;
  return hipSuccess;
}
}  // namespace

template <typename R>
hipError_t prepare_workspace(int ndim, int A, int H, int W, int rytov, size_t *bytes) {
  PrepLayout L;
  hipError_t e = prep_layout<R>(ndim, A, H, W, rytov != 0, &L);
  *bytes = L.total;
  return e;
}

template <typename R>
hipError_t launch_prepare(int ndim, int A, int H, int W, const void *u, const void *u0,
                          int rytov, int align, void *out, void *ws, size_t ws_bytes,
                          hipStream_t st) {
  PrepLayout L;
  hipError_t e = prep_layout<R>(ndim, A, H, W, rytov != 0, &L);
  if (e != hipSuccess) return e;
  if (ws_bytes < L.total) return hipErrorInvalidValue;
  char *w = (char *)ws;
  const long long n = (long long)A * H * W;
  R *phase = (R *)(w + L.phase);
  R *lna = rytov ? (R *)(w + L.lna) : nullptr;
  k_prep_phase<R><<<blocks_of(n), 256, 0, st>>>((const cpx<R> *)u, (const cpx<R> *)u0, n,
                                                 phase, lna);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ndim == 2) {
    k_unwrap_rows<R><<<(A + 63) / 64, 64, 0, st>>>(phase, A, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  } else {
    const long long E = (long long)H * (W - 1) + (long long)(H - 1) * W;
    double *rel = (double *)(w + L.rel);
    k_uw_reliability<R><<<blocks_of(n), 256, 0, st>>>(phase, A, H, W, rel);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    double *k0 = (double *)(w + L.keys0), *k1 = (double *)(w + L.keys1);
    int *v0 = (int *)(w + L.vals0), *v1 = (int *)(w + L.vals1);
    unsigned *offs = (unsigned *)(w + L.offs);
    const long long ne = std::max<long long>((long long)A * E, (long long)A + 1);
    k_uw_edges<R><<<blocks_of(ne), 256, 0, st>>>(rel, phase, A, H, W, k0, v0, offs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sb = L.sort_bytes;
    e = rocprim::segmented_radix_sort_pairs(w + L.sort_tmp, sb, k0, k1, v0, v1,
                                            (unsigned)(A * E), (unsigned)A, offs, offs + 1,
                                            0, 64, st);
    if (e != hipSuccess) return e;
    int2 *node = (int2 *)(w + L.parent);
    // one image per block: images advance independently.  Images of at
    // most 65536 pixels keep their group sizes in LDS (ODTBRAIN_AMD_UW_GLOBAL
    // set: the global union-find for every size)
    if ((long long)H * W <= kUwLdsPix && !std::getenv("ODTBRAIN_AMD_UW_GLOBAL")) {
      const size_t sh = uw_claims_offset((long long)H * W) +
                        (kUwClaims + 2 * 5 * 64) * sizeof(unsigned);
      if (sh > 65536 &&
          (e = hipFuncSetAttribute((const void *)k_uw_merge_lds,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh)) !=
              hipSuccess)
        return e;
      k_uw_merge_lds<<<A, 128, sh, st>>>(A, H, W, v1, node);
    } else {
      k_uw_gather<R><<<A, kUwThreads, 0, st>>>(A, H, W, v1, node);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_uw_apply<R><<<blocks_of(n), 256, 0, st>>>(phase, A, H, W, node);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  double *remove = nullptr;
  if (align) {
    R *part = (R *)(w + L.part);
    const int nparts = (int)std::min<long long>(kMinParts, (n + 255) / 256);
    k_min_blocks<R><<<nparts, 256, 0, st>>>(phase, n, part);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    remove = (double *)(w + L.remove);
    k_align<R><<<1, 256, 0, st>>>(phase, A, ndim, H, W, part, nparts, (R *)(w + L.samp),
                                  (double *)(w + L.steps), remove);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  k_finish<R><<<blocks_of(n), 256, 0, st>>>(phase, lna, (const cpx<R> *)u0, remove, n,
                                            (long long)H * W, out);
  return hipGetLastError();
}

template hipError_t prepare_workspace<float>(int, int, int, int, int, size_t *);
template hipError_t prepare_workspace<double>(int, int, int, int, int, size_t *);
template hipError_t launch_prepare<float>(int, int, int, int, const void *, const void *,
                                          int, int, void *, void *, size_t, hipStream_t);
template hipError_t launch_prepare<double>(int, int, int, int, const void *, const void *,
                                           int, int, void *, void *, size_t, hipStream_t);

}  // namespace odt

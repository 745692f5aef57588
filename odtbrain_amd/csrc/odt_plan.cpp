// Plan construction and the C ABI of libodtbrain_hip.so (include/odtbrain_hip.h).
//
// A plan holds the float64-derived tables of one reconstruction call of the
// reference (odtbrain/_alg3d_bpp.py:371-441, _alg3d_bppt.py:736-822,
// _alg2d_bpp.py:239-287): twiddles, the low-pass support box of the
// spectrum, the prefactor table, the propagator phase table km*(M-1), the
// fused B-spline prefilter terms, and the per-angle weights and geometry.
// Execution streams the projections through the transform kernels in
// chunks and accumulates batches of coefficient slabs into the output.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/odtbrain_hip.h"
#include "odt_launch.h"

using odt::cpx;
using odt::Geom;
using odt::PropAux;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Fault injection for tests of the error paths, on the host: with
// ODTBRAIN_AMD_FAULT="alloc:N" the N-th device allocation the library makes
// after the variable took that value reports hipErrorOutOfMemory instead of
// allocating; with "launch:N" the N-th checked launch or HIP call reports
// hipErrorLaunchFailure after it was issued (so the kernels themselves run
// and end normally).  Each fault fires once per value of the variable.
std::mutex g_fault_mu;
std::string g_fault_spec;
long long g_fault_count[2] = {0, 0};
long long g_live_allocs = 0;  // device allocations of the library not yet freed

bool fault_fires(int kind) {  // 0: allocation, 1: launch
  const char *env = std::getenv("ODTBRAIN_AMD_FAULT");
  if (!env || !*env) return false;
  std::lock_guard<std::mutex> lk(g_fault_mu);
  if (g_fault_spec != env) {
    g_fault_spec = env;
    g_fault_count[0] = g_fault_count[1] = 0;
  }
  const char *what = kind == 0 ? "alloc:" : "launch:";
  const size_t wl = std::strlen(what);
  if (g_fault_spec.compare(0, wl, what) != 0) return false;
  const long long n = std::atoll(g_fault_spec.c_str() + wl);
  return ++g_fault_count[kind] == n;
}

// ROCTx ranges around the stages a call enqueues (ODTBRAIN_AMD_ROCTX=1:
// rocprofv3 --marker-trace shows them beside the kernels).  They mark the
// host-side enqueue of each stage, the work itself runs asynchronously on
// the stream.  The roctx library is opened on first use: no link-time
// dependency, and nothing is called when the variable is unset.
struct RoctxApi {
  int (*push)(const char *) = nullptr;
  int (*pop)() = nullptr;
};
const RoctxApi &roctx_api() {
  static const RoctxApi api = [] {
    RoctxApi a;
    const char *env = std::getenv("ODTBRAIN_AMD_ROCTX");
    if (!env || env[0] == '\0' || env[0] == '0') return a;
    void *h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_GLOBAL);
    if (!h) h = dlopen("libroctx64.so.4", RTLD_NOW | RTLD_GLOBAL);
    if (!h) return a;
    a.push = reinterpret_cast<int (*)(const char *)>(dlsym(h, "roctxRangePushA"));
    a.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
    if (!a.push || !a.pop) a = RoctxApi{};
    return a;
  }();
  return api;
}
struct StageRange {
  bool on;
  explicit StageRange(const char *name) : on(roctx_api().push != nullptr) {
    if (on) roctx_api().push(name);
  }
  ~StageRange() {
    if (on) roctx_api().pop();
  }
  StageRange(const StageRange &) = delete;
  StageRange &operator=(const StageRange &) = delete;
};

hipError_t dev_malloc(void **ptr, size_t bytes) {
  if (fault_fires(0)) {
    *ptr = nullptr;
    return hipErrorOutOfMemory;
  }
  const hipError_t e = hipMalloc(ptr, bytes);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_fault_mu);
    ++g_live_allocs;
  }
  return e;
}
template <typename T>
hipError_t dev_malloc(T **ptr, size_t bytes) {
  return dev_malloc(reinterpret_cast<void **>(ptr), bytes);
}

void dev_free(void *ptr) {
  if (!ptr) return;
  if (hipFree(ptr) == hipSuccess) {
    std::lock_guard<std::mutex> lk(g_fault_mu);
    --g_live_allocs;
  }
}

hipError_t checked(hipError_t e) {
  return e == hipSuccess && fault_fires(1) ? hipErrorLaunchFailure : e;
}

int hip_fail(hipError_t e, const char *where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return ODT_ERR_HIP;
}

#define HIPCHK(expr, where)                             \
  do {                                                  \
    hipError_t e_ = checked(expr);                      \
    if (e_ != hipSuccess) return hip_fail(e_, where);   \
  } while (0)

int ilog2_exact(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Device allocation released at scope exit (hipFree waits for the device).
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) dev_free(p);
  }
};

// B-spline samples at 0 and +-1 (beta(theta) = b0 + 2 b1 cos theta).
void bspline_b01(int order, double *b0, double *b1) {
  if (order == 2) {
    *b0 = 0.75;
    *b1 = 0.125;
  } else {
    *b0 = 2.0 / 3.0;
    *b1 = 1.0 / 6.0;
  }
}

// First column of the inverse mirror-boundary B-spline sampling matrix:
// the recursive prefilter (exact mirror initialisation) applied to delta_0.
std::vector<double> prefilter_delta0(int order, int n) {
  std::vector<double> c(n, 0.0);
  c[0] = 1.0;
  if (n < 2) return c;
  const double z = order == 2 ? std::sqrt(8.0) - 3.0 : std::sqrt(3.0) - 2.0;
  const double gain = (1.0 - z) * (1.0 - 1.0 / z);
  for (auto &v : c) v *= gain;
  const double zn1 = std::pow(z, n - 1);
  double c0 = c[0] + zn1 * c[n - 1];
  double zi = z;
  for (int i = 1; i < n - 1; ++i) {
    c0 += zi * (c[i] + zn1 * c[n - 1 - i]);
    zi *= z;
  }
  c[0] = c0 / (1.0 - zn1 * zn1);
  for (int i = 1; i < n; ++i) c[i] += z * c[i - 1];
  c[n - 1] = (z * c[n - 2] + c[n - 1]) * z / (z * z - 1.0);
  for (int i = n - 2; i >= 0; --i) c[i] = z * (c[i + 1] - c[i]);
  return c;
}

// Frequency support along one axis: signed indices j (numpy.fft.fftfreq
// order, j in [-(n/2), (n-1)/2]) with (2*pi*j/n)^2 < km^2.
void support(int n, double km, int *j0, int *len) {
  const int jmin = -(n / 2), jmax = (n - 1) / 2;
  const double val = 1.0 / (double)n;
  const double two_pi = 2.0 * M_PI;
  int J = -1;
  for (int j = 0; j <= jmax || j <= -jmin; ++j) {
    const double k = two_pi * ((double)j * val);
    if (k * k < km * km)
      J = j;
    else
      break;
  }
  if (J < 0) {
    *j0 = 0;
    *len = 1;
    return;
  }
  const int lo = std::max(jmin, -J), hi = std::min(jmax, J);
  *j0 = lo;
  *len = hi - lo + 1;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct odt_plan {
  int kind, precision, order, onlyreal, flags, pstride, A;
  int device = 0;  // HIP device the plan's tables live on
  double padval_re, padval_im;
  Geom g;
  bool explicit_iir;
  int NB, NC;                 // slabs per accumulate batch, angles per chunk
  int RG;                     // angles per rotate-accumulate launch
  int ldt;                    // row stride of T (support columns, padded)
  int nb_fr, nb_fc, nb_col, nb_row;  // transforms per block
  size_t esz_c, esz_v;        // bytes of complex element / slab element
  size_t off_F1, off_S, off_T, off_slab, off_comp, off_win, off_part, ws_bytes;
  size_t off_gt = 0;  // adjoint: the output gradient, planes innermost (or 0)
  // Column and row stages of consecutive chunks overlapped (opt-in): the
  // row stage of chunk c runs on a second stream while the column stage of
  // chunk c + 1 fills the other of two T buffers (off_T2).
  bool overlap = false;
  size_t off_T2 = 0;
  hipStream_t side = nullptr;
  hipEvent_t ev_col[2] = {nullptr, nullptr}, ev_row[2] = {nullptr, nullptr};
  int RGG = 1;  // concurrent angle groups of the 2D rotate-accumulate
  // 2D: the row stage forms the column stage's product on load (no T)
  bool fuse2d = false;
  // forward stage in float64 (always for FP64 plans; for FP32 plans when
  // the transforms fit the LDS in float64); the float64 twiddles and
  // prefactor it uses (the same tables as twx/twy/P in FP64 plans)
  bool fwd64 = true;
  void *twx64 = nullptr, *twy64 = nullptr, *P64 = nullptr;
  // device tables
  void *twx = nullptr, *twy = nullptr, *P = nullptr, *aux = nullptr;
  void *q0x = nullptr, *q0y = nullptr, *q0z = nullptr;
  double *prop_a = nullptr, *weights = nullptr, *params = nullptr;
  void *gtab = nullptr;      // propagator table of the column stage (or null)
  void *fir_taps = nullptr;  // adjoint: prefilter impulse response
  int fir_K = 0;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_rot;  // start/stop pairs
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
  int n_rot = 0;    // timed rotate stages (one per slab batch)
  int n_rot_k = 0;  // rotate-accumulate kernel launches of the last execute
};

namespace {

// The per-plan propagator table of the column stage fits (2 GiB at most).
bool prop_table_fits(const Geom &g, size_t esz) {
  const size_t tab = (size_t)g.ln * g.Lx * g.Ly * esz;
  return tab <= ((size_t)2 << 30) && !std::getenv("ODTBRAIN_AMD_NO_PROP_TABLE");
}

template <typename R>
int upload_tables(odt_plan *p, const odt_desc *d) {
  const Geom &g = p->g;
  const int Lx = g.Lx, Ly = g.Ly;
  const double km = d->km, lD = d->lD;
  const double two_pi = 2.0 * M_PI;
  const double dphi0 = two_pi / (double)d->A;
  double b0 = 0, b1 = 0;
  const bool fused = !p->explicit_iir && p->order >= 2;
  if (fused) bspline_b01(p->order, &b0, &b1);
  p->g.b1 = b1;
  const std::complex<double> pre0 =
      std::complex<double>(0.0, -1.0) * km / two_pi * dphi0;
  const double norm = 1.0 / ((double)g.lNx * (double)g.lNy);
  std::vector<cpx<R>> P((size_t)Lx * Ly);
  const bool sep64 = p->fwd64 && sizeof(R) != sizeof(double);
  std::vector<cpx<double>> P64(sep64 ? (size_t)Lx * Ly : 0);
  std::vector<double> pa((size_t)Lx * Ly);
  std::vector<PropAux<R>> aux(g.fuse_z ? (size_t)Lx * Ly : 1);
  const double valx = 1.0 / (double)g.lNx, valy = 1.0 / (double)g.lNy;
  for (int c = 0; c < Lx; ++c) {
    const int jx = g.jx0 + c;
    const double kx = two_pi * ((double)jx * valx);
    const double bx = g.fuse_x ? b0 + 2.0 * b1 * std::cos(two_pi * jx * valx) : 1.0;
    for (int cy = 0; cy < Ly; ++cy) {
      const int jy = g.jy0 + cy;
      const double ky = two_pi * ((double)jy * valy);
      const double klp = (kx * kx + ky * ky < km * km) ? 1.0 : 0.0;
      const double M = 1.0 / km * std::sqrt((km * km - kx * kx - ky * ky) * klp);
      const double ramp = std::fabs(kx * d->ramp_v + ky * d->ramp_u) * klp;
      std::complex<double> pf = pre0 * ramp;
      pf *= std::exp(std::complex<double>(0.0, -km * (M - 1.0) * lD));
      pf *= norm;
      if (g.fuse_x) pf /= bx;
      if (g.fuse_y) pf /= b0 + 2.0 * b1 * std::cos(two_pi * jy * valy);
      const size_t t = (size_t)c * Ly + cy;
      P[t] = odt::mk<R>((R)pf.real(), (R)pf.imag());
      if (sep64) P64[t] = odt::mk<double>(pf.real(), pf.imag());
      const double a = km * (M - 1.0);
      pa[t] = a;
      if (g.fuse_z) {
        const double beta = b0 + 2.0 * b1 * std::cos(a);
        auto h = [&](double k) {
          return std::exp(std::complex<double>(0.0, a * (g.z0 + k))) / beta;
        };
        const std::complex<double> r0 = b1 * (h(-1.0) - h(1.0));
        const std::complex<double> r1 = b1 * (h((double)g.ln) - h((double)g.ln - 2.0));
        aux[t].r0 = odt::mk<R>((R)r0.real(), (R)r0.imag());
        aux[t].r1 = odt::mk<R>((R)r1.real(), (R)r1.imag());
        aux[t].ib = (R)(1.0 / beta);
        aux[t].pad = 0;
      }
    }
  }
  auto q0vec = [&](int n, bool on) {
    std::vector<R> q(std::max(n, 1), (R)0);
    if (on) {
      // entries below 1e-20 are far under either precision's rounding of
      // the O(1) terms they multiply; zero lets kernels skip them
      auto qd = prefilter_delta0(p->order, n);
      for (int i = 0; i < n; ++i) q[i] = std::fabs(qd[i]) < 1e-20 ? (R)0 : (R)qd[i];
    }
    return q;
  };
  const auto qx = q0vec(g.Nx, g.fuse_x), qy = q0vec(g.Ny, g.fuse_y),
             qz = q0vec(g.ln, g.fuse_z);
  HIPCHK(dev_malloc(&p->P, P.size() * sizeof(cpx<R>)), "hipMalloc(P)");
  HIPCHK(hipMemcpy(p->P, P.data(), P.size() * sizeof(cpx<R>), hipMemcpyHostToDevice), "copy P");
  HIPCHK(dev_malloc((void **)&p->prop_a, pa.size() * sizeof(double)), "hipMalloc(prop)");
  HIPCHK(hipMemcpy(p->prop_a, pa.data(), pa.size() * sizeof(double), hipMemcpyHostToDevice), "copy prop");
  HIPCHK(dev_malloc(&p->aux, aux.size() * sizeof(PropAux<R>)), "hipMalloc(aux)");
  HIPCHK(hipMemcpy(p->aux, aux.data(), aux.size() * sizeof(PropAux<R>), hipMemcpyHostToDevice), "copy aux");
  HIPCHK(dev_malloc(&p->q0x, qx.size() * sizeof(R)), "hipMalloc(q0x)");
  HIPCHK(hipMemcpy(p->q0x, qx.data(), qx.size() * sizeof(R), hipMemcpyHostToDevice), "copy q0x");
  HIPCHK(dev_malloc(&p->q0y, qy.size() * sizeof(R)), "hipMalloc(q0y)");
  HIPCHK(hipMemcpy(p->q0y, qy.data(), qy.size() * sizeof(R), hipMemcpyHostToDevice), "copy q0y");
  HIPCHK(dev_malloc(&p->q0z, qz.size() * sizeof(R)), "hipMalloc(q0z)");
  HIPCHK(hipMemcpy(p->q0z, qz.data(), qz.size() * sizeof(R), hipMemcpyHostToDevice), "copy q0z");
  // transform tables (Bluestein tables for lengths that are not powers of
  // two, built in float64 and rounded once for float32 plans)
  const size_t ntx = (size_t)odt::dft_table_len(g.lNx), nty = (size_t)odt::dft_table_len(g.lNy);
  DevBuf scratch64;
  if (sizeof(R) != sizeof(double) &&
      (odt::uses_bluestein(g.lNx) || odt::uses_bluestein(g.lNy)))
    HIPCHK(dev_malloc(&scratch64.p, std::max(ntx, nty) * sizeof(cpx<double>)),
           "hipMalloc(table scratch)");
  cpx<double> *s64 = (cpx<double> *)scratch64.p;
  HIPCHK(dev_malloc(&p->twx, ntx * sizeof(cpx<R>)), "hipMalloc(twx)");
  HIPCHK(dev_malloc(&p->twy, nty * sizeof(cpx<R>)), "hipMalloc(twy)");
  HIPCHK(odt::launch_dft_table<R>((cpx<R> *)p->twx, g.lNx, s64, 0), "twiddles x");
  if (s64) HIPCHK(hipDeviceSynchronize(), "twiddles x");
  HIPCHK(odt::launch_dft_table<R>((cpx<R> *)p->twy, g.lNy, s64, 0), "twiddles y");
  if (sep64) {
    HIPCHK(dev_malloc(&p->P64, P64.size() * sizeof(cpx<double>)), "hipMalloc(P64)");
    HIPCHK(hipMemcpy(p->P64, P64.data(), P64.size() * sizeof(cpx<double>),
                     hipMemcpyHostToDevice), "copy P64");
    HIPCHK(dev_malloc(&p->twx64, ntx * sizeof(cpx<double>)), "hipMalloc(twx64)");
    HIPCHK(dev_malloc(&p->twy64, nty * sizeof(cpx<double>)), "hipMalloc(twy64)");
    HIPCHK(odt::launch_dft_table<double>((cpx<double> *)p->twx64, g.lNx, nullptr, 0),
           "twiddles x64");
    HIPCHK(odt::launch_dft_table<double>((cpx<double> *)p->twy64, g.lNy, nullptr, 0),
           "twiddles y64");
  } else if (p->fwd64) {
    p->twx64 = p->twx;
    p->twy64 = p->twy;
    p->P64 = p->P;
  }
  // The column stage's propagator (times the fused z-prefilter terms) is
  // the same for every angle: tabulated once per plan when it fits in
  // 2 GiB (C2: 256 x 227 x 227 complex64, 106 MB).
  if (prop_table_fits(g, sizeof(cpx<R>))) {
    HIPCHK(dev_malloc(&p->gtab, (size_t)g.ln * g.Lx * g.Ly * sizeof(cpx<R>)),
           "hipMalloc(gtab)");
    HIPCHK(odt::launch_prop_table<R>(p->prop_a, (const PropAux<R> *)p->aux,
                                     (const R *)p->q0z, g, (cpx<R> *)p->gtab, 0),
           "k_prop_table");
  }
  HIPCHK(hipDeviceSynchronize(), "plan tables");
  return ODT_OK;
}

int per_block(int n, size_t esz, int cap) {
  size_t b = 65536 / ((size_t)2 * odt::lds_len(n) * esz);
  if (b < 1) b = 1;
  if (b > 16) b = 16;
  return std::max(1, std::min((int)b, cap));
}

void free_plan(odt_plan *p) {
  void *ptrs[] = {p->twx, p->twy, p->P, p->aux, p->q0x, p->q0y, p->q0z,
                  p->prop_a, p->weights, p->params, p->fir_taps, p->gtab};
  for (void *q : ptrs)
    if (q) dev_free(q);
  if (p->P64 && p->P64 != p->P) dev_free(p->P64);
  if (p->twx64 && p->twx64 != p->twx) dev_free(p->twx64);
  if (p->twy64 && p->twy64 != p->twy) dev_free(p->twy64);
  for (auto e : p->ev_rot) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    if (p->ev_col[i]) (void)hipEventDestroy(p->ev_col[i]);
    if (p->ev_row[i]) (void)hipEventDestroy(p->ev_row[i]);
  }
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->ev_t0) (void)hipEventDestroy(p->ev_t0);
  if (p->ev_t1) (void)hipEventDestroy(p->ev_t1);
  delete p;
}

template <typename R, typename V>
int execute_t(odt_plan *p, const void *d_usin, void *d_out, char *ws,
              hipStream_t st) {
  const Geom &g = p->g;
  cpx<R> *F1 = (cpx<R> *)(ws + p->off_F1);
  cpx<R> *S = (cpx<R> *)(ws + p->off_S);
  cpx<R> *T = (cpx<R> *)(ws + p->off_T);
  V *slab = (V *)(ws + p->off_slab);
  V *comp = p->off_comp ? (V *)(ws + p->off_comp) : nullptr;
  V *out = (V *)d_out;
  const cpx<R> *usin = (const cpx<R> *)d_usin;
  const bool tilted = p->kind == ODT_KIND_3D_TILTED;
  StageRange call_range("odt.execute");
  if (p->timing) HIPCHK(hipEventRecord(p->ev_t0, st), "event");
  p->n_rot = 0;
  p->n_rot_k = 0;
  // two streams only outside graph capture (a captured batch keeps one
  // stream per plan)
  bool overlap = p->overlap;
  if (overlap) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
    overlap = cs == hipStreamCaptureStatusNone;
  }
  // row stages a failed earlier call left on the side stream finish before
  // this call reuses the T buffers (waits on unrecorded events are no-ops)
  if (overlap)
    for (int i = 0; i < 2; ++i) HIPCHK(hipStreamWaitEvent(st, p->ev_row[i], 0), "event wait");
  int ci = 0;  // chunk counter over the whole call
  for (int b0 = 0; b0 < p->A; b0 += p->NB) {
    const int nb = std::min(p->NB, p->A - b0);
    for (int c0 = b0; c0 < b0 + nb; c0 += p->NC, ++ci) {
      const int n = std::min(p->NC, b0 + nb - c0);
      const int tb = overlap ? ci & 1 : 0;
      cpx<R> *Tc = tb ? (cpx<R> *)(ws + p->off_T2) : T;
      hipStream_t rst = overlap ? p->side : st;
      StageRange chunk_range("odt.filter_chunk");
      if (p->fwd64) {
        cpx<double> *F1d = (cpx<double> *)(ws + p->off_F1);
        HIPCHK((odt::launch_fwd_rows<R, double>(usin, p->weights, F1d,
                                                (const cpx<double> *)p->twx64, g,
                                                p->padval_re, p->padval_im, c0, n,
                                                p->nb_fr, st)),
               "k_fwd_rows");
        HIPCHK((odt::launch_fwd_cols<R, double>(F1d, S, (const cpx<double> *)p->twy64,
                                                (const cpx<double> *)p->P64, g, n,
                                                p->nb_fc, st)),
               "k_fwd_cols");
      } else {
        HIPCHK((odt::launch_fwd_rows<R, R>(usin, p->weights, F1, (const cpx<R> *)p->twx, g,
                                           p->padval_re, p->padval_im, c0, n, p->nb_fr,
                                           st)),
               "k_fwd_rows");
        HIPCHK((odt::launch_fwd_cols<R, R>(F1, S, (const cpx<R> *)p->twy,
                                           (const cpx<R> *)p->P, g, n, p->nb_fc, st)),
               "k_fwd_cols");
      }
      // the row stage of chunk ci - 2 has read this T buffer
      if (overlap && ci >= 2) HIPCHK(hipStreamWaitEvent(st, p->ev_row[tb], 0), "event wait");
      if (!p->fuse2d)
        HIPCHK(odt::launch_col_ifft<R>(S, Tc, (const cpx<R> *)p->twy, p->prop_a,
                                       (const PropAux<R> *)p->aux, (const R *)p->q0z,
                                       (const R *)p->q0y, g, n, p->nb_col, p->ldt,
                                       (const cpx<R> *)p->gtab, st),
               "k_col_ifft");
      // (the side stream's wait covers everything queued on st before, the
      // previous batch's rotate-accumulate included)
      if (overlap) {
        HIPCHK(hipEventRecord(p->ev_col[tb], st), "event");
        HIPCHK(hipStreamWaitEvent(rst, p->ev_col[tb], 0), "event wait");
      }
      HIPCHK((odt::launch_row_ifft<R, V>(
                 Tc, slab, (const cpx<R> *)p->twx, (const R *)p->q0x, g, n, p->nb_row,
                 c0 - b0, p->ldt, rst, p->fuse2d ? S : nullptr,
                 p->fuse2d ? (const cpx<R> *)p->gtab : nullptr)),
             "k_row_ifft");
      if (overlap) HIPCHK(hipEventRecord(p->ev_row[tb], rst), "event");
    }
    // every row stage of the batch has written its slabs (the side stream
    // runs them in order)
    if (overlap) HIPCHK(hipStreamWaitEvent(st, p->ev_row[(ci - 1) & 1], 0), "event wait");
    if (p->explicit_iir) {
      const long long ec = sizeof(V) / sizeof(R);
      const long long Nx = g.Nx, Ny = g.Ny, ln = g.ln;
      // along x
      HIPCHK(odt::launch_iir<R>((R *)slab, nb * Ny * ln * ec, g.Nx, ec, 1, Nx * ec, ec,
                                p->order, st), "iir x");
      // along z
      HIPCHK(odt::launch_iir<R>((R *)slab, nb * Ny * Nx * ec, g.ln, Nx * ec, 1,
                                ln * Nx * ec, Nx * ec, p->order, st), "iir z");
      if (tilted)
        HIPCHK(odt::launch_iir<R>((R *)slab, nb * ln * Nx * ec, g.Ny, ln * Nx * ec, 1,
                                  Ny * ln * Nx * ec, ln * Nx * ec, p->order, st),
               "iir y");
    }
    const int first = b0 == 0, last = b0 + nb == p->A;
    const double *prm = p->params + (size_t)b0 * p->pstride;
    if (p->timing) {
      if ((int)p->ev_rot.size() < 2 * (p->n_rot + 1)) {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0), "event");
        HIPCHK(hipEventCreate(&e1), "event");
        p->ev_rot.push_back(e0);
        p->ev_rot.push_back(e1);
      }
      HIPCHK(hipEventRecord(p->ev_rot[2 * p->n_rot], st), "event");
    }
    p->n_rot_k += !tilted && p->order < 2 ? 1 : (nb + p->RG - 1) / p->RG;
    StageRange rot_range(tilted ? "odt.affine_accumulate" : "odt.rotate_accumulate");
    if (tilted)
      for (int g0 = 0; g0 < nb; g0 += p->RG) {
        const int gn = std::min(p->RG, nb - g0);
        const size_t per_angle = (size_t)g.Ny * g.ln * g.Nx;
        HIPCHK((odt::launch_affine_acc<R, V>(p->order, slab + (size_t)g0 * per_angle, gn,
                                             prm + (size_t)g0 * p->pstride,
                                             (int *)(ws + p->off_win), out, comp,
                                             g.Ny, g.ln, g.Nx, first && g0 == 0,
                                             last && g0 + gn == nb, st)),
               "k_affine_acc");
      }
    else if (p->kind == ODT_KIND_3D && p->order < 2)
      HIPCHK((odt::launch_seq_rotate_sum<R, V>(p->order, slab, nb, prm, out, comp,
                                               g.Ny, g.ln, g.Nx, first, last, st)),
             "k_seq_rotate");
    else
      for (int g0 = 0; g0 < nb; g0 += p->RG) {
        const int gn = std::min(p->RG, nb - g0);
        const size_t per_angle = (size_t)g.Ny * g.ln * g.Nx;
        HIPCHK((odt::launch_rotate_acc<R, V>(
                   p->order, slab + (size_t)g0 * per_angle, gn,
                   prm + (size_t)g0 * p->pstride, (int *)(ws + p->off_win), out,
                   comp, g.Ny, g.ln, g.Nx, first && g0 == 0,
                   last && g0 + gn == nb, st, p->RGG,
                   p->off_part ? (V *)(ws + p->off_part) : nullptr)),
               "k_rotate_acc");
      }
    if (p->timing) HIPCHK(hipEventRecord(p->ev_rot[2 * p->n_rot + 1], st), "event");
    p->n_rot++;
  }
  if (p->timing) HIPCHK(hipEventRecord(p->ev_t1, st), "event");
  return ODT_OK;
}

// Two-sided impulse response of the mirror-boundary B-spline prefilter of
// an order: per pole z the factor (1 - z)^2 / ((1 - z q)(1 - z / q)) has
// the response (1 - z) / (1 + z) z^|k|; the poles' responses are convolved
// and truncated where they fall below 1e-18 of the centre tap.  h[k + K].
std::vector<double> bspline_fir(int order, int *K) {
  int np = 0;
  double z1 = 0, z2 = 0, gain = 1;
  odt::iir_poles(order, &np, &z1, &z2, &gain);
  std::vector<double> h(1, 1.0);
  for (int pi = 0; pi < np; ++pi) {
    const double z = pi == 0 ? z1 : z2;
    const int kp = (int)std::ceil(std::log(1e-18) / std::log(std::fabs(z)));
    std::vector<double> hp(2 * kp + 1);
    for (int k = -kp; k <= kp; ++k)
      hp[k + kp] = (1 - z) / (1 + z) * std::pow(z, (double)std::abs(k));
    std::vector<double> c(h.size() + hp.size() - 1, 0.0);
    for (size_t i = 0; i < h.size(); ++i)
      for (size_t j = 0; j < hp.size(); ++j) c[i + j] += h[i] * hp[j];
    h.swap(c);
  }
  const int kc = (int)(h.size() - 1) / 2;
  int kk = 0;
  for (int k = 0; k <= kc; ++k)
    if (std::fabs(h[kc + k]) > 1e-18 * std::fabs(h[kc]) ||
        std::fabs(h[kc - k]) > 1e-18 * std::fabs(h[kc]))
      kk = k;
  *K = kk;
  return std::vector<double>(h.begin() + (kc - kk), h.begin() + (kc + kk + 1));
}

// Prefilter impulse response of the adjoint's transposed prefilter, uploaded
// at plan creation (so that odt_plan_adjoint only launches kernels).
template <typename R>
int upload_fir(odt_plan *p) {
  int K = 0;
  const std::vector<double> h = bspline_fir(p->order, &K);
  const int pad = odt::adj_fir_pad();
  std::vector<R> hz(2 * K + 2 * pad, R(0));
  for (int k = -K; k <= K; ++k) hz[k + K + pad] = (R)h[k + K];
  HIPCHK(dev_malloc(&p->fir_taps, hz.size() * sizeof(R)), "hipMalloc(taps)");
  HIPCHK(hipMemcpy(p->fir_taps, hz.data(), hz.size() * sizeof(R), hipMemcpyHostToDevice),
         "hipMemcpy(taps)");
  p->fir_K = K;
  return ODT_OK;
}

// Adjoint of execute_t for the y-axis reconstruction (see odt_adjoint.hip).
template <typename R>
int adjoint_t(odt_plan *p, const void *d_gout, void *d_gusin, char *ws,
              hipStream_t st) {
  StageRange range("odt.adjoint");
  const Geom &g = p->g;
  cpx<R> *U1 = (cpx<R> *)(ws + p->off_F1);
  cpx<R> *Sp = (cpx<R> *)(ws + p->off_S);
  cpx<R> *Tp = (cpx<R> *)(ws + p->off_T);
  cpx<R> *slab = (cpx<R> *)(ws + p->off_slab);
  const cpx<R> *gout = (const cpx<R> *)d_gout;
  cpx<R> *gu = (cpx<R> *)d_gusin;
  const long long Nx = g.Nx, Ny = g.Ny, ln = g.ln;
  const size_t per_angle = (size_t)Ny * ln * Nx;
  // Fused plans (orders 2, 3) transpose the prefilter inside the
  // transforms (k_adj_row's boundary terms, the table's conjugate in
  // k_adj_col).  Otherwise, with room for two half batches the transposed
  // prefilter runs as two out-of-place convolutions (z into the second
  // half, x back), else as the in-place recursions.
  const bool fused = !p->explicit_iir;
  // The gather reads the gradient transposed to planes-innermost
  // (k_adj_rotate_t) when the plan has room for the copy, else as it is
  // (k_adj_rotate).
  const bool tr = p->off_gt != 0 && !std::getenv("ODTBRAIN_AMD_ADJ_DIRECT");
  const bool fir = !fused && p->NB >= 2;
  const int step = fir ? p->NB / 2 : p->NB;
  cpx<R> *slabB = slab + (size_t)step * per_angle;
  cpx<R> *gt = tr ? (cpx<R> *)(ws + p->off_gt) : nullptr;
  if (fir && !p->fir_taps)
    return fail(ODT_ERR_UNSUPPORTED, "adjoint: the plan has no prefilter taps");
  if (tr) HIPCHK(odt::launch_adj_transpose<R>(gout, gt, g.Ny, g.ln, g.Nx, st), "k_adj_transpose");
  for (int b0 = 0; b0 < p->A; b0 += step) {
    const int nb = std::min(step, p->A - b0);
    const double *prm = p->params + 6 * (size_t)b0;
    HIPCHK(tr ? odt::launch_adj_rotate_t<R>(p->order, gt, nb, prm, slab, g.Ny, g.ln, g.Nx, st)
              : odt::launch_adj_rotate<R>(p->order, gout, nb, prm, slab, g.Ny, g.ln, g.Nx,
                                          st),
           "k_adj_rotate");
    if (fused) {
      // nothing: the transforms below carry the transposed prefilter
    } else if (fir) {
      const R *hz = (const R *)p->fir_taps;
      HIPCHK(odt::launch_adj_fir<R>(slab, slabB, nb * Ny * Nx, g.ln, Nx, 1, ln * Nx, Nx,
                                    hz, p->fir_K, st),
             "k_adj_fir z");
      HIPCHK(odt::launch_adj_fir_rows<R>(slabB, slab, nb * Ny * ln, g.Nx, hz, p->fir_K, st),
             "k_adj_fir_rows");
    } else {
    // Q^T = D Q D^-1 along x, then along z (real and imaginary lines)
    R *sr = (R *)slab;
    const long long lz = nb * Ny * Nx * 2;
    HIPCHK(odt::launch_adj_prefilter_rows<R>(slab, nb * Ny * ln, g.Nx, p->order, st),
           "k_adj_iir_rows");
    HIPCHK(odt::launch_scale_ends<R>(sr, lz, g.ln, Nx * 2, 1, ln * Nx * 2, Nx * 2, 2.0, st),
           "ends z");
    HIPCHK(odt::launch_iir<R>(sr, lz, g.ln, Nx * 2, 1, ln * Nx * 2, Nx * 2, p->order, st),
           "iir z");
    HIPCHK(odt::launch_scale_ends<R>(sr, lz, g.ln, Nx * 2, 1, ln * Nx * 2, Nx * 2, 0.5, st),
           "ends z");
    }
    for (int c0 = b0; c0 < b0 + nb; c0 += p->NC) {
      const int n = std::min(p->NC, b0 + nb - c0);
      HIPCHK(odt::launch_adj_fft<R>(0, g.lNx, slab + (size_t)(c0 - b0) * per_angle, Tp,
                                    p->twx, g.fuse_x ? p->q0x : nullptr, nullptr, nullptr,
                                    0, g, n, p->ldt, st),
             "k_adj_row");
      HIPCHK(odt::launch_adj_fft<R>(1, g.lNy, Tp, Sp, p->twy, p->gtab, p->prop_a, nullptr,
                                    0, g, n, p->ldt, st),
             "k_adj_col");
      HIPCHK(odt::launch_adj_fft<R>(2, g.lNy, Sp, U1, p->twy, p->P, nullptr, nullptr, 0, g,
                                    n, g.Lx, st),
             "k_adj_inv_col");
      HIPCHK(odt::launch_adj_fft<R>(3, g.lNx, U1, gu, p->twx, nullptr, nullptr, p->weights,
                                    c0, g, n, g.Lx, st),
             "k_adj_inv_row");
    }
  }
  return ODT_OK;
}

// Workspace of the allocation-free entry point (odt_plan_execute_noalloc):
// a file-scope device array, one instance per GPU the library's code object
// is loaded on, held until the process ends.  Host code never uses the
// array's name as an address: its device address is looked up once per
// device with hipGetSymbolAddress and kept in scratch_ptr[].
static __device__ __align__(256) unsigned char scratch[ODT_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `dev` (the current device), looked up on
// the first call for that device.
int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(ODT_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(ODT_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    scratch_ptr[dev] = ptr;
  }
  *out = scratch_ptr[dev];
  return ODT_OK;
}

}  // namespace

extern "C" {

const char *odt_last_error(void) { return g_err.c_str(); }

int odt_debug_live_allocations(int64_t *count) {
  if (!count) return fail(ODT_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(g_fault_mu);
  *count = g_live_allocs;
  return ODT_OK;
}

int odt_version(void) { return 1; }

int odt_plan_create(const odt_desc *d, odt_plan **out) {
  return odt_plan_create_tuned(d, nullptr, out);
}

int odt_plan_create_tuned(const odt_desc *d, const odt_tuning *tu, odt_plan **out) {
  if (!d || !out) return fail(ODT_ERR_INVALID, "null argument");
  if (tu && (tu->rotate_group < 0 || tu->rotate_groups_2d < 0 || tu->t_budget_bytes < 0 ||
             tu->col_per_angle < 0 || tu->col_per_angle > 2))
    return fail(ODT_ERR_INVALID, "tuning: negative or unknown values");
  *out = nullptr;
  if (d->kind < 0 || d->kind > 2) return fail(ODT_ERR_INVALID, "unknown kind");
  if (d->precision != ODT_FP32 && d->precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision must be ODT_FP32 or ODT_FP64");
  if (d->A < 1 || d->Ny < 1 || d->Nx < 1)
    return fail(ODT_ERR_INVALID, "A, Ny, Nx must be positive");
  if (d->lNy < d->Ny || d->lNx < d->Nx || d->padyl < 0 || d->padxl < 0 ||
      d->padyl + d->Ny > d->lNy || d->padxl + d->Nx > d->lNx)
    return fail(ODT_ERR_INVALID, "inconsistent padding");
  if (d->pad_mode < 0 || d->pad_mode > 2)
    return fail(ODT_ERR_INVALID, "pad_mode");
  if (d->pad_mode == ODT_PAD_NONE && (d->lNy != d->Ny || d->lNx != d->Nx))
    return fail(ODT_ERR_INVALID, "pad_mode NONE needs lNy == Ny and lNx == Nx");
  if (d->intp_order < 0 || d->intp_order > 5)
    return fail(ODT_ERR_INVALID, "intp_order must be in 0..5");
  const int want_stride = d->kind == ODT_KIND_3D_TILTED ? 12 : 6;
  if (d->angle_param_stride != want_stride)
    return fail(ODT_ERR_INVALID, "angle_param_stride must be 6 (planar) or 12 (affine)");
  if (!d->weights || !d->angle_params)
    return fail(ODT_ERR_INVALID, "weights and angle_params are required");
  if (d->kind == ODT_KIND_3D_TILTED) {
    // the LDS windows of the affine gather are sized for rotations: a row of
    // a rotation matrix has an L1 norm of at most sqrt(3)
    for (int a = 0; a < d->A; ++a)
      for (int r = 0; r < 3; ++r) {
        const double *m = d->angle_params + 12 * (size_t)a + 3 * r;
        if (!(std::fabs(m[0]) + std::fabs(m[1]) + std::fabs(m[2]) <= std::sqrt(3.0) + 1e-9))
          return fail(ODT_ERR_INVALID, "angle_params: each 3x3 matrix must be a rotation");
      }
  }
  if (!(d->km > 0)) return fail(ODT_ERR_INVALID, "km must be positive");
  if (d->kind == ODT_KIND_2D && (d->Ny != 1 || d->lNy != 1))
    return fail(ODT_ERR_INVALID, "2D plans need Ny == lNy == 1");
  const size_t esz = d->precision == ODT_FP64 ? 16 : 8;
  const size_t max_lds = 160 * 1024;
  const size_t llx = (size_t)odt::lds_len(d->lNx), lly = (size_t)odt::lds_len(d->lNy);
  if (2 * llx * esz > max_lds || 2 * lly * esz > max_lds)
    return fail(ODT_ERR_UNSUPPORTED, "transform length exceeds the LDS capacity");

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  odt_plan *p = new odt_plan();
  p->device = dev;
  p->kind = d->kind;
  p->precision = d->precision;
  p->order = d->intp_order;
  p->onlyreal = d->onlyreal ? 1 : 0;
  p->flags = d->flags;
  p->pstride = want_stride;
  p->A = d->A;
  p->padval_re = d->padval_re;
  p->padval_im = d->padval_im;
  Geom &g = p->g;
  std::memset(&g, 0, sizeof(g));
  g.A = d->A;
  g.Ny = d->Ny;
  g.Nx = d->Nx;
  g.ln = d->Nx;
  g.lNy = d->lNy;
  g.lNx = d->lNx;
  g.logNy = ilog2_exact(d->lNy);
  g.logNx = ilog2_exact(d->lNx);
  g.padyl = d->padyl;
  g.padxl = d->padxl;
  g.pad_mode = d->pad_mode;
  g.z0 = d->z0;
  support(g.lNx, d->km, &g.jx0, &g.Lx);
  support(g.lNy, d->km, &g.jy0, &g.Ly);
  const bool pf = p->order >= 2;
  const bool want_iir = (d->flags & ODT_FLAG_IIR_PREFILTER) != 0;
  const bool fusable = p->order == 2 || p->order == 3;
  p->explicit_iir = pf && (want_iir || !fusable);
  const bool fused = pf && !p->explicit_iir;
  g.fuse_x = fused && g.Nx > 1;
  g.fuse_z = fused && g.ln > 1;
  g.fuse_y = fused && d->kind == ODT_KIND_3D_TILTED && g.Ny > 1;
  g.col_mode = tu ? tu->col_per_angle : 0;

  p->esz_c = esz;
  p->esz_v = p->onlyreal ? esz / 2 : esz;
  const size_t slab_angle = (size_t)g.Ny * g.ln * g.Nx * p->esz_v;
  size_t budget = d->slab_budget_bytes > 0 ? (size_t)d->slab_budget_bytes
                                           : ((size_t)8 << 30);
  p->NB = (int)std::max<size_t>(1, std::min<size_t>(budget / slab_angle, (size_t)d->A));
  p->ldt = (g.Lx + 15) & ~15;
  const size_t t_angle = (size_t)g.ln * p->ldt * g.Ny * esz;
  // Angles per filter chunk: the intermediate T of a chunk is written by the
  // column stage and read back by the row stage.  Larger chunks measured
  // faster at 256^2 x 360 (T budget 1 GiB: 90.8 ms per reconstruction,
  // 4 GiB: 87.4, 8 GiB: 86.9; smaller than 1 GiB: slower), so up to 8 GiB,
  // at most a 32nd of the device memory.  Tuning knob (bytes).
  const char *tenv = std::getenv("ODTBRAIN_AMD_T_BUDGET");
  size_t t_budget = (size_t)8 << 30;
  {
    // from the total memory (not the free memory): identical descriptions
    // give identical plans
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
      t_budget = std::min(t_budget, std::max(total_b / 32, (size_t)1 << 30));
  }
  if (tenv) t_budget = (size_t)std::atoll(tenv);
  if (tu && tu->t_budget_bytes > 0) t_budget = (size_t)tu->t_budget_bytes;
  const size_t nc_budget = std::max<size_t>(1, t_budget / t_angle);
  p->NC = (int)std::min<size_t>(nc_budget, (size_t)p->NB);
  p->NC = std::min(p->NC, 64);
  // float64 forward stage where its transforms fit the LDS in float64
  p->fwd64 = d->precision == ODT_FP64 ||
             (2 * llx * 16 <= max_lds && 2 * lly * 16 <= max_lds);
  const size_t esz_f = p->fwd64 ? 16 : esz;
  p->nb_fr = per_block(g.lNx, esz_f, g.lNy);
  p->nb_fc = per_block(g.lNy, esz_f, g.Lx);
  p->nb_col = per_block(g.lNy, esz, g.ln);
  // rows of the row stage: the slices when there is one detector row
  p->nb_row = per_block(g.lNx, esz, g.Ny == 1 ? std::max(1, g.ln) : g.Ny);

  p->fuse2d = d->kind == ODT_KIND_2D && prop_table_fits(g, esz) &&
              (d->precision == ODT_FP64 ? odt::row_fuses_2d<double>(g, true)
                                        : odt::row_fuses_2d<float>(g, true)) &&
              !std::getenv("ODTBRAIN_AMD_NO_FUSE_2D");
  // Optionally (ODTBRAIN_AMD_OVERLAP=1), when a batch takes two chunks or
  // more, the column and row stages of consecutive chunks overlap through
  // two T buffers, which share the T budget when it binds.  It paid while
  // the row stage was latency-bound (256^2 x 360 complex64: 57.8 -> 56.1 ms
  // per reconstruction); with both stages near the copy rate it costs
  // (48.9 -> 50.0 ms), so one stream is the default.  Fused 2D plans have
  // no T: one chunk per slab batch (F1 and S are a row per angle).
  if (p->fuse2d) p->NC = p->NB;
  const char *ovenv = std::getenv("ODTBRAIN_AMD_OVERLAP");
  p->overlap = !p->fuse2d && p->NC < p->NB && ovenv && *ovenv && *ovenv != '0';
  if (p->overlap) p->NC = std::max(1, std::min(p->NC, (int)std::min<size_t>(nc_budget / 2, 64)));

  size_t off = 0;
  p->off_F1 = off;
  off = align256(off + (size_t)p->NC * g.Lx * g.lNy * esz_f);
  p->off_S = off;
  off = align256(off + (size_t)p->NC * g.Lx * g.Ly * esz);
  p->off_T = off;
  if (!p->fuse2d) off = align256(off + (size_t)p->NC * g.ln * p->ldt * g.Ny * esz);
  if (p->overlap) {
    p->off_T2 = off;
    off = align256(off + (size_t)p->NC * g.ln * p->ldt * g.Ny * esz);
  }
  p->off_slab = off;
  off = align256(off + (size_t)p->NB * slab_angle);
  // Angles per rotate-accumulate launch.  A block keeps its output tile in
  // registers across the angles of one launch; blocks that share input
  // planes drift apart over many angles and then no longer share them in
  // L2, so long batches are split (the running sums go through the output
  // and its compensation term between launches).  Measured at 256^2 x 360
  // (complex64, four-plane rotate blocks, Z-ordered tiles): 36 angles 14.6
  // ms per reconstruction, 48: 14.4, 72: 14.5, 120: 14.9; the tilted gather
  // keeps the 48 it was measured with.
  p->RG = p->NB;
  if ((d->kind == ODT_KIND_3D || d->kind == ODT_KIND_3D_TILTED) && d->intp_order >= 2) {
    const char *env = std::getenv("ODTBRAIN_AMD_ROTATE_GROUP");
    const int rg = tu && tu->rotate_group > 0 ? tu->rotate_group : env ? std::atoi(env) : 48;
    if (rg > 0) p->RG = std::min(p->NB, rg);
  }
  p->off_comp = 0;
  if (p->NB < d->A || p->RG < p->NB) {
    p->off_comp = off;
    off = align256(off + slab_angle);
  }
  // windows of the rotate-accumulate tiles (four ints per tile/angle): the
  // planar tiles of a batch, or the 3D tiles of one tilted launch
  p->off_win = off;
  off = align256(off + 4 * sizeof(int) *
                           (d->kind == ODT_KIND_3D_TILTED
                                ? (size_t)p->RG * odt::affine_window_entries(
                                                      p->g.Ny, p->g.ln, p->g.Nx)
                                : (size_t)p->NB *
                                      odt::rotate_window_entries(p->g.ln, p->g.Nx)));
  // 2D images have few tiles ((N/16)^2: 64 at N = 128) and each block loops
  // over every angle, so the angles run in concurrent groups (at least 8
  // angles each, about 2560 blocks in all, a sum and a compensation volume
  // per group) that are added up in group order afterwards.
  p->RGG = 1;
  p->off_part = 0;
  if (d->kind == ODT_KIND_2D) {
    const size_t tiles = odt::rotate_window_entries(p->g.ln, p->g.Nx);
    int gg = (int)((2560 + tiles - 1) / tiles);
    gg = std::min(gg, std::max(1, p->NB / 8));
    if (const char *genv = std::getenv("ODTBRAIN_AMD_ROTATE_2D_GROUPS"))
      gg = std::max(1, std::atoi(genv));
    if (tu && tu->rotate_groups_2d > 0) gg = tu->rotate_groups_2d;
    if (gg > 1) {
      p->RGG = gg;
      p->off_part = off;
      off = align256(off + 2 * (size_t)gg * slab_angle);
    }
  }
  // the adjoint's copy of the output gradient with the planes innermost
  // (one volume; plans whose adjoint can use it)
  p->off_gt = 0;
  if (d->kind == ODT_KIND_3D && !p->onlyreal && p->order >= 2 &&
      (d->precision == ODT_FP64
           ? odt::adj_rotate_t_supported<double>(p->order, g.Ny, g.ln, g.Nx)
           : odt::adj_rotate_t_supported<float>(p->order, g.Ny, g.ln, g.Nx))) {
    p->off_gt = off;
    off = align256(off + slab_angle);
  }
  p->ws_bytes = off;

  int rc;
  hipError_t e;
  if ((e = dev_malloc((void **)&p->weights, sizeof(double) * d->A)) != hipSuccess ||
      (e = hipMemcpy(p->weights, d->weights, sizeof(double) * d->A,
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = dev_malloc((void **)&p->params, sizeof(double) * d->A * want_stride)) !=
          hipSuccess ||
      (e = hipMemcpy(p->params, d->angle_params, sizeof(double) * d->A * want_stride,
                     hipMemcpyHostToDevice)) != hipSuccess) {
    free_plan(p);
    return hip_fail(e, "plan upload");
  }
  rc = d->precision == ODT_FP64 ? upload_tables<double>(p, d) : upload_tables<float>(p, d);
  // plans the adjoint supports get its prefilter taps now
  if (rc == ODT_OK && p->kind == ODT_KIND_3D && !p->onlyreal && p->explicit_iir &&
      p->order >= 2)
    rc = d->precision == ODT_FP64 ? upload_fir<double>(p) : upload_fir<float>(p);
  if (rc != ODT_OK) {
    std::string keep = g_err;
    free_plan(p);
    g_err = keep;
    return rc;
  }
  if ((e = hipEventCreate(&p->ev_t0)) != hipSuccess ||
      (e = hipEventCreate(&p->ev_t1)) != hipSuccess) {
    free_plan(p);
    return hip_fail(e, "event");
  }
  if (p->overlap) {
    for (int i = 0; i < 2 && e == hipSuccess; ++i)
      if ((e = checked(hipEventCreateWithFlags(&p->ev_col[i], hipEventDisableTiming))) ==
          hipSuccess)
        e = checked(hipEventCreateWithFlags(&p->ev_row[i], hipEventDisableTiming));
    if (e == hipSuccess) e = checked(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    if (e != hipSuccess) {
      free_plan(p);
      return hip_fail(e, "overlap stream");
    }
  }
  *out = p;
  return ODT_OK;
}

int odt_plan_workspace_size(const odt_plan *p, size_t *bytes) {
  if (!p || !bytes) return fail(ODT_ERR_INVALID, "null argument");
  *bytes = p->ws_bytes;
  return ODT_OK;
}

int odt_plan_execute(odt_plan *p, const void *d_usin, void *d_out, void *ws,
                     size_t ws_bytes, void *stream) {
  if (!p || !d_usin || !d_out) return fail(ODT_ERR_INVALID, "null argument");
  if (ws_bytes < p->ws_bytes || (!ws && p->ws_bytes))
    return fail(ODT_ERR_WORKSPACE, "workspace too small: need " +
                                       std::to_string(p->ws_bytes) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  char *w = (char *)ws;
  if (p->precision == ODT_FP64)
    return p->onlyreal ? execute_t<double, double>(p, d_usin, d_out, w, st)
                       : execute_t<double, cpx<double>>(p, d_usin, d_out, w, st);
  return p->onlyreal ? execute_t<float, float>(p, d_usin, d_out, w, st)
                     : execute_t<float, cpx<float>>(p, d_usin, d_out, w, st);
}

int odt_plan_execute_noalloc(odt_plan *p, const void *d_usin, void *d_out,
                             void *stream) {
  if (!p || !d_usin || !d_out) return fail(ODT_ERR_INVALID, "null argument");
  if (p->timing)
    return fail(ODT_ERR_INVALID, "odt_plan_execute_noalloc: timing enabled on the plan");
  if (p->ws_bytes > (size_t)ODT_SCRATCH_BYTES)
    return fail(ODT_ERR_WORKSPACE,
                "odt_plan_execute_noalloc: the plan needs " + std::to_string(p->ws_bytes) +
                    " bytes of workspace, the static scratch holds " +
                    std::to_string((size_t)ODT_SCRATCH_BYTES) +
                    " (create the plan with a smaller slab_budget_bytes)");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev), "hipGetDevice");
  if (dev != p->device)
    return fail(ODT_ERR_INVALID, "odt_plan_execute_noalloc: the current device is not "
                                 "the plan's device");
  void *ws = nullptr;
  const int rc = scratch_address(dev, &ws);
  if (rc != ODT_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  char *w = (char *)ws;
  if (p->precision == ODT_FP64)
    return p->onlyreal ? execute_t<double, double>(p, d_usin, d_out, w, st)
                       : execute_t<double, cpx<double>>(p, d_usin, d_out, w, st);
  return p->onlyreal ? execute_t<float, float>(p, d_usin, d_out, w, st)
                     : execute_t<float, cpx<float>>(p, d_usin, d_out, w, st);
}

int odt_plan_destroy(odt_plan *p) {
  if (p) free_plan(p);
  return ODT_OK;
}

int odt_plan_adjoint(odt_plan *p, const void *d_gout, void *d_gusin, void *ws,
                     size_t ws_bytes, void *stream) {
  if (!p || !d_gout || !d_gusin) return fail(ODT_ERR_INVALID, "null argument");
  if (p->kind != ODT_KIND_3D || p->onlyreal || p->order < 2 ||
      (p->g.pad_mode != ODT_PAD_NONE && p->g.pad_mode != ODT_PAD_EDGE))
    return fail(ODT_ERR_UNSUPPORTED,
                "adjoint: y-axis 3D plans with complex output, spline order "
                "2..5 and edge or no padding only");
  if (!p->explicit_iir && (!p->gtab || p->g.fuse_y))
    return fail(ODT_ERR_UNSUPPORTED,
                "adjoint: fused-prefilter plans need their propagator table");
  auto pow2_ok = [](int n) { return n == 64 || n == 128 || n == 256 || n == 512; };
  if (!pow2_ok(p->g.lNx) || !pow2_ok(p->g.lNy))
    return fail(ODT_ERR_UNSUPPORTED, "adjoint: padded sizes 64..512 (powers of two) only");
  if (p->g.Nx < 8)
    return fail(ODT_ERR_UNSUPPORTED, "adjoint: detector width of at least 8 pixels");
  if (ws_bytes < p->ws_bytes || !ws)
    return fail(ODT_ERR_WORKSPACE, "workspace too small: need " +
                                       std::to_string(p->ws_bytes) + " bytes");
  hipStream_t st = (hipStream_t)stream;
  return p->precision == ODT_FP64
             ? adjoint_t<double>(p, d_gout, d_gusin, (char *)ws, st)
             : adjoint_t<float>(p, d_gout, d_gusin, (char *)ws, st);
}

struct odt_batch {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int n = 0;
  std::vector<hipStream_t> lanes;  // capture streams (default priority)
  std::vector<hipEvent_t> events;  // fork and joins
};

namespace {

void free_batch(odt_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto s : b->lanes) (void)hipStreamDestroy(s);
  for (auto e : b->events) (void)hipEventDestroy(e);
  delete b;
}

// Estimated work of a plan (filter transforms plus resampling), for
// balancing the lanes of a batch.
double plan_cost(const odt_plan *p) {
  const Geom &g = p->g;
  return (double)p->A * g.ln * ((double)g.lNy * g.lNx + (double)g.Ny * g.Nx);
}

}  // namespace

int odt_batch_create_lanes(int n, odt_plan *const *plans, const void *const *d_usin,
                           void *const *d_out, void *ws, size_t ws_bytes, int lanes,
                           void *stream, odt_batch **out) {
  if (!out || n < 1 || !plans || !d_usin || !d_out || !stream || lanes < 1 || !ws)
    return fail(ODT_ERR_INVALID, "bad arguments (a non-null stream is required)");
  *out = nullptr;
  lanes = std::min(lanes, n);
  // each lane owns an equal, 256-byte aligned part of the workspace
  const size_t part = (ws_bytes / (size_t)lanes) & ~(size_t)255;
  for (int i = 0; i < n; ++i) {
    if (!plans[i] || !d_usin[i] || !d_out[i])
      return fail(ODT_ERR_INVALID, "null plan or buffer in batch");
    if (plans[i]->timing) return fail(ODT_ERR_INVALID, "timing enabled in a batch plan");
    if (part < plans[i]->ws_bytes)
      return fail(ODT_ERR_WORKSPACE, "workspace too small for batch entry " +
                                         std::to_string(i) + " (need " +
                                         std::to_string(plans[i]->ws_bytes) +
                                         " bytes per lane)");
  }
  // longest first onto the least loaded lane
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return plan_cost(plans[a]) > plan_cost(plans[b]);
  });
  std::vector<int> lane_of(n);
  std::vector<double> load(lanes, 0.0);
  for (int i : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lane_of[i] = l;
    load[l] += plan_cost(plans[i]);
  }
  auto lane_ws = [&](int l) { return (void *)((char *)ws + (size_t)l * part); };
  hipStream_t st = (hipStream_t)stream;
  // one eager pass: validates every entry and sets kernel attributes
  // outside the capture
  for (int i = 0; i < n; ++i) {
    const int rc = odt_plan_execute(plans[i], d_usin[i], d_out[i], lane_ws(lane_of[i]),
                                    part, stream);
    if (rc != ODT_OK) return rc;
  }
  HIPCHK(hipStreamSynchronize(st), "batch warm-up");
  odt_batch *b = new odt_batch;
  b->n = n;
  hipError_t e = hipSuccess;
  for (int l = 0; l < lanes && e == hipSuccess; ++l) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) b->lanes.push_back(s);
  }
  for (int k = 0; k < lanes + 1 && e == hipSuccess; ++k) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) b->events.push_back(ev);
  }
  if (e != hipSuccess) {
    free_batch(b);
    return hip_fail(e, "batch streams");
  }
  int rc = ODT_OK;
  e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) {
    hipError_t ec = hipEventRecord(b->events[0], st);  // fork
    for (int l = 0; l < lanes && ec == hipSuccess; ++l)
      ec = hipStreamWaitEvent(b->lanes[l], b->events[0], 0);
    for (int i : order) {
      if (ec != hipSuccess || rc != ODT_OK) break;
      const int l = lane_of[i];
      rc = odt_plan_execute(plans[i], d_usin[i], d_out[i], lane_ws(l), part, b->lanes[l]);
    }
    for (int l = 0; l < lanes && ec == hipSuccess; ++l) {  // joins
      ec = hipEventRecord(b->events[l + 1], b->lanes[l]);
      if (ec == hipSuccess) ec = hipStreamWaitEvent(st, b->events[l + 1], 0);
    }
    const hipError_t e2 = hipStreamEndCapture(st, &b->graph);
    e = ec != hipSuccess ? ec : e2;
  }
  if (e == hipSuccess && rc == ODT_OK)
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess || rc != ODT_OK) {
    const std::string keep = g_err;
    free_batch(b);
    if (rc != ODT_OK) {
      g_err = keep;
      return rc;
    }
    return hip_fail(e, "batch capture");
  }
  *out = b;
  return ODT_OK;
}

int odt_batch_create(int n, odt_plan *const *plans, const void *const *d_usin,
                     void *const *d_out, void *ws, size_t ws_bytes,
                     void *stream, odt_batch **out) {
  return odt_batch_create_lanes(n, plans, d_usin, d_out, ws, ws_bytes, 1, stream, out);
}

int odt_batch_launch(odt_batch *b, void *stream) {
  if (!b) return fail(ODT_ERR_INVALID, "null batch");
  StageRange range("odt.batch_launch");
  HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
  return ODT_OK;
}

int odt_batch_destroy(odt_batch *b) {
  if (b) free_batch(b);
  return ODT_OK;
}

int odt_plan_enable_timing(odt_plan *p, int enable) {
  if (!p) return fail(ODT_ERR_INVALID, "null plan");
  p->timing = enable != 0;
  return ODT_OK;
}

int odt_plan_last_timing(const odt_plan *p, double *rot_ms, double *filt_ms,
                         double *total_ms, int *launches) {
  if (!p) return fail(ODT_ERR_INVALID, "null plan");
  if (!p->timing) return fail(ODT_ERR_INVALID, "timing not enabled");
  HIPCHK(hipEventSynchronize(p->ev_t1), "event sync");
  float t = 0;
  double rot = 0;
  for (int i = 0; i < p->n_rot; ++i) {
    HIPCHK(hipEventElapsedTime(&t, p->ev_rot[2 * i], p->ev_rot[2 * i + 1]), "elapsed");
    rot += t;
  }
  HIPCHK(hipEventElapsedTime(&t, p->ev_t0, p->ev_t1), "elapsed");
  if (rot_ms) *rot_ms = rot;
  if (total_ms) *total_ms = t;
  if (filt_ms) *filt_ms = t - rot;
  if (launches) *launches = p->n_rot_k;
  return ODT_OK;
}

int odt_fft_c2c(int precision, int64_t batch, int32_t n, int32_t direction,
                const void *d_in, void *d_out, void *stream) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (batch < 0 || n < 1 || (direction != 1 && direction != -1) || !d_in || !d_out)
    return fail(ODT_ERR_INVALID, "bad fft arguments");
  const size_t esz = precision == ODT_FP64 ? 16 : 8;
  if (2 * (size_t)odt::lds_len(n) * esz > 160 * 1024)
    return fail(ODT_ERR_UNSUPPORTED, "transform length exceeds the LDS capacity");
  if (batch == 0) return ODT_OK;
  hipStream_t st = (hipStream_t)stream;
  const size_t ntw = (size_t)odt::dft_table_len(n);
  DevBuf tw, scratch64;
  HIPCHK(dev_malloc(&tw.p, ntw * esz), "hipMalloc(tw)");
  if (precision == ODT_FP32 && odt::uses_bluestein(n))
    HIPCHK(dev_malloc(&scratch64.p, ntw * 16), "hipMalloc(table scratch)");
  const int nb = per_block(n, esz, (int)std::min<int64_t>(batch, 16));
  const int logn = ilog2_exact(n);
  hipError_t e;
  if (precision == ODT_FP64) {
    e = odt::launch_dft_table<double>((cpx<double> *)tw.p, n, nullptr, st);
    if (e == hipSuccess)
      e = odt::launch_fft_batch<double>((const cpx<double> *)d_in, (cpx<double> *)d_out,
                                        batch, n, logn, direction,
                                        (const cpx<double> *)tw.p, nb, st);
  } else {
    e = odt::launch_dft_table<float>((cpx<float> *)tw.p, n, (cpx<double> *)scratch64.p, st);
    if (e == hipSuccess)
      e = odt::launch_fft_batch<float>((const cpx<float> *)d_in, (cpx<float> *)d_out,
                                       batch, n, logn, direction,
                                       (const cpx<float> *)tw.p, nb, st);
  }
  hipError_t e2 = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "k_fft_batch");
  if (e2 != hipSuccess) return hip_fail(e2, "k_fft_batch sync");
  return ODT_OK;
}

int odt_odt_to_ri(int precision, int is_complex, int64_t n, double res, double nm,
                  const void *d_f, void *d_ri, void *stream) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (n < 0 || (n > 0 && (!d_f || !d_ri)) || !(res != 0.0))
    return fail(ODT_ERR_INVALID, "bad arguments");
  const double km = (2 * M_PI * nm) / res;
  const double km2 = km * km;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = precision == ODT_FP64
                     ? odt::launch_odt_to_ri<double>(is_complex, n, km2, nm, d_f, d_ri, st)
                     : odt::launch_odt_to_ri<float>(is_complex, n, km2, nm, d_f, d_ri, st);
  if (e != hipSuccess) return hip_fail(e, "k_odt_to_ri");
  return ODT_OK;
}

int odt_opt_to_ri(int precision, int is_complex, int64_t n, double res, double nm,
                  const void *d_f, void *d_ri, void *stream) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (n < 0 || (n > 0 && (!d_f || !d_ri))) return fail(ODT_ERR_INVALID, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = precision == ODT_FP64
                     ? odt::launch_opt_to_ri<double>(is_complex, n, res, nm, d_f, d_ri, st)
                     : odt::launch_opt_to_ri<float>(is_complex, n, res, nm, d_f, d_ri, st);
  if (e != hipSuccess) return hip_fail(e, "k_opt_to_ri");
  return ODT_OK;
}

namespace {
int prepare_check(int precision, int ndim, int32_t A, int32_t Ny, int32_t Nx) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (ndim != 2 && ndim != 3) return fail(ODT_ERR_INVALID, "ndim must be 2 or 3");
  if (A < 1 || Nx < 3 || (ndim == 3 && Ny < 2))
    return fail(ODT_ERR_INVALID, "sinogram too small");
  const long long n = (long long)A * (ndim == 3 ? Ny : 1) * Nx;
  const long long E = ndim == 3 ? 2LL * Ny * Nx : 0;
  if (n > (1LL << 31) - 1 || (long long)A * E > 0xffffffffLL)
    return fail(ODT_ERR_UNSUPPORTED, "sinogram too large");
  return ODT_OK;
}
}  // namespace

int odt_prepare_workspace_size(int precision, int ndim, int32_t A, int32_t Ny,
                               int32_t Nx, int rytov, size_t *bytes) {
  if (!bytes) return fail(ODT_ERR_INVALID, "null argument");
  const int rc = prepare_check(precision, ndim, A, Ny, Nx);
  if (rc != ODT_OK) return rc;
  const int H = ndim == 3 ? Ny : 1;
  hipError_t e = precision == ODT_FP64
                     ? odt::prepare_workspace<double>(ndim, A, H, Nx, rytov, bytes)
                     : odt::prepare_workspace<float>(ndim, A, H, Nx, rytov, bytes);
  if (e != hipSuccess) return hip_fail(e, "prepare_workspace");
  return ODT_OK;
}

int odt_sinogram_prepare(int precision, int ndim, int32_t A, int32_t Ny, int32_t Nx,
                         const void *d_usin, const void *d_u0, int rytov, int align,
                         void *d_out, void *d_ws, size_t ws_bytes, void *stream) {
  const int rc = prepare_check(precision, ndim, A, Ny, Nx);
  if (rc != ODT_OK) return rc;
  if (!d_usin || !d_out || !d_ws) return fail(ODT_ERR_INVALID, "null argument");
  StageRange range("odt.sinogram_prepare");
  size_t need = 0;
  const int H = ndim == 3 ? Ny : 1;
  HIPCHK(precision == ODT_FP64 ? odt::prepare_workspace<double>(ndim, A, H, Nx, rytov, &need)
                               : odt::prepare_workspace<float>(ndim, A, H, Nx, rytov, &need),
         "prepare_workspace");
  if (ws_bytes < need)
    return fail(ODT_ERR_WORKSPACE, "workspace too small: need " + std::to_string(need) +
                                       " bytes");
  hipStream_t st = (hipStream_t)stream;
  hipError_t e =
      precision == ODT_FP64
          ? odt::launch_prepare<double>(ndim, A, H, Nx, d_usin, d_u0, rytov, align, d_out,
                                        d_ws, ws_bytes, st)
          : odt::launch_prepare<float>(ndim, A, H, Nx, d_usin, d_u0, rytov, align, d_out,
                                       d_ws, ws_bytes, st);
  if (e != hipSuccess) return hip_fail(e, "odt_sinogram_prepare");
  return ODT_OK;
}

namespace {
struct AppleLayout {
  size_t data, ft, ft0, tmp, env, core, hp, hp2, cp, part, sum, nnan, tw0, tw1, tw2, w,
      shpart, mm, hist, total;
  int r;
};

AppleLayout apple_layout(int n0, int n1, int n2) {
  AppleLayout L{};
  const size_t N = (size_t)n0 * n1 * n2;
  const int m = std::max(n0, n1);
  const size_t M = (size_t)m * m * m;
  size_t o = 0;
  auto take = [&](size_t b) {
    const size_t at = o;
    o = align256(o + b);
    return at;
  };
  L.data = take(N * 16);
  L.ft = take(N * 16);
  L.ft0 = take(N * 16);
  L.tmp = take(N * 16);
  L.env = take(N * 8);
  L.core = take(N);
  L.hp = take(M * 8);
  L.hp2 = take(M * 8);
  L.cp = take(M);
  L.part = take(((N + 255) / 256) * 8);
  L.sum = take(8);
  L.nnan = take(4);
  L.tw0 = take((size_t)odt::dft_table_len(n0) * 16);
  L.tw1 = take((size_t)odt::dft_table_len(n1) * 16);
  L.tw2 = take((size_t)odt::dft_table_len(n2) * 16);
  const double sigma = (double)std::max(n0, std::max(n1, n2)) / 100.0;
  L.r = (int)(4.0 * sigma + 0.5);
  L.w = take((size_t)(2 * L.r + 1) * 8);
  L.shpart = take(odt::sh_part_bytes());
  L.mm = take(8 * 8);
  L.hist = take(odt::sh_hist_bytes());
  L.total = o;
  return L;
}

// 3D transform of a (n0, n1, n2) complex128 array in place: batched 1D
// transforms along the contiguous axis, the other two axes brought there
// by batched transposes.
hipError_t fft3(cpx<double> *a, cpx<double> *tmp, int n0, int n1, int n2, int dir,
                const cpx<double> *tw0, const cpx<double> *tw1, const cpx<double> *tw2,
                hipStream_t st) {
  hipError_t e;
  auto fft = [&](cpx<double> *x, long long batch, int n, const cpx<double> *tw) {
    return odt::launch_fft_batch<double>(x, x, batch, n, ilog2_exact(n), dir, tw,
                                         per_block(n, 16, 16), st);
  };
  if ((e = fft(a, (long long)n0 * n1, n2, tw2)) != hipSuccess) return e;
  if ((e = odt::launch_transpose3(a, tmp, n0, n1, n2, st)) != hipSuccess) return e;
  if ((e = fft(tmp, (long long)n0 * n2, n1, tw1)) != hipSuccess) return e;
  if ((e = odt::launch_transpose3(tmp, a, n0, n2, n1, st)) != hipSuccess) return e;
  if ((e = odt::launch_transpose3(a, tmp, 1, n0, n1 * n2, st)) != hipSuccess) return e;
  if ((e = fft(tmp, (long long)n1 * n2, n0, tw0)) != hipSuccess) return e;
  return odt::launch_transpose3(tmp, a, 1, n1 * n2, n0, st);
}
}  // namespace

int odt_apple_workspace_size(int32_t n0, int32_t n1, int32_t n2, size_t *bytes) {
  if (!bytes || n0 < 1 || n1 < 1 || n2 < 1) return fail(ODT_ERR_INVALID, "bad arguments");
  *bytes = apple_layout(n0, n1, n2).total;
  return ODT_OK;
}

int odt_apple_core(int32_t n0, int32_t n1, int32_t n2, double res, double nm,
                   uint8_t *d_core, void *stream) {
  if (n0 < 1 || n1 < 1 || n2 < 1 || !d_core) return fail(ODT_ERR_INVALID, "bad arguments");
  if (n0 != n2) return fail(ODT_ERR_INVALID, "`shape[0]` and `shape[2]` must be identical!");
  const double km = (2 * M_PI * nm) / res;
  HIPCHK(odt::launch_apple_core(d_core, n0, n1, n2, km, (hipStream_t)stream), "k_apple_core");
  return ODT_OK;
}

int odt_apple_correct(int f_precision, int32_t n0, int32_t n1, int32_t n2, const void *d_f,
                      double res, double nm, int method, const uint8_t *d_mask,
                      double bg_shell_width, double enforce_envelope, int32_t max_iter,
                      double min_diff, int out_precision, void *d_out, void *d_ws,
                      size_t ws_bytes, void *stream, int32_t *iterations) {
  if ((f_precision != ODT_FP32 && f_precision != ODT_FP64) ||
      (out_precision != ODT_FP32 && out_precision != ODT_FP64))
    return fail(ODT_ERR_INVALID, "precision");
  if (n0 < 1 || n1 < 1 || n2 < 1 || !d_f || !d_out || !d_ws || max_iter < 1)
    return fail(ODT_ERR_INVALID, "bad arguments");
  if (n0 != n2) return fail(ODT_ERR_INVALID, "`shape[0]` and `shape[2]` must be identical!");
  if (!(enforce_envelope >= 0 && enforce_envelope <= 1))
    return fail(ODT_ERR_INVALID, "`enforce_envelope` must be in interval [0, 1]");
  if (method != 0 && method != 1)
    return fail(ODT_ERR_INVALID, "method must be 0 (non-negativity) or 1 (symmetric histogram)");
  const int mmax = std::max(n0, std::max(n1, n2));
  const int lmax = std::max(odt::lds_len(n0), std::max(odt::lds_len(n1), odt::lds_len(n2)));
  if (2 * (size_t)lmax * 16 > 160 * 1024 || n0 > 65535)
    return fail(ODT_ERR_UNSUPPORTED, "volume edge too long for the transforms");
  const AppleLayout L = apple_layout(n0, n1, n2);
  if (ws_bytes < L.total)
    return fail(ODT_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) +
                                       " bytes");
  hipStream_t st = (hipStream_t)stream;
  char *w = (char *)d_ws;
  cpx<double> *data = (cpx<double> *)(w + L.data), *ft = (cpx<double> *)(w + L.ft),
              *ft0 = (cpx<double> *)(w + L.ft0), *tmp = (cpx<double> *)(w + L.tmp);
  double *env = (double *)(w + L.env);
  unsigned char *core = (unsigned char *)(w + L.core);
  cpx<double> *tw0 = (cpx<double> *)(w + L.tw0), *tw1 = (cpx<double> *)(w + L.tw1),
              *tw2 = (cpx<double> *)(w + L.tw2);
  const long long N = (long long)n0 * n1 * n2;
  HIPCHK(odt::launch_dft_table<double>(tw0, n0, nullptr, st), "twiddles");
  HIPCHK(odt::launch_dft_table<double>(tw1, n1, nullptr, st), "twiddles");
  HIPCHK(odt::launch_dft_table<double>(tw2, n2, nullptr, st), "twiddles");
  const double km = (2 * M_PI * nm) / res;
  HIPCHK(odt::launch_apple_core(core, n0, n1, n2, km, st), "k_apple_core");
  HIPCHK(odt::launch_apple_init(d_f, f_precision == ODT_FP64, data, N, st), "k_apple_init");
  HIPCHK(hipMemcpyAsync(ft0, data, N * 16, hipMemcpyDeviceToDevice, st), "copy");
  HIPCHK(fft3(ft0, tmp, n0, n1, n2, -1, tw0, tw1, tw2, st), "fft3");
  const bool use_env = enforce_envelope != 0.0;
  if (use_env) {
    // scipy.ndimage gaussian weights, computed on the stream
    HIPCHK(odt::launch_gauss_weights((double *)(w + L.w), L.r, (double)mmax / 100.0, st),
           "k_gauss_weights");
    HIPCHK(hipMemsetAsync(w + L.nnan, 0, 4, st), "memset");
    double *g = (double *)tmp, *g2 = (double *)tmp + N;
    HIPCHK(odt::launch_envelope(ft0, core, n0, n1, n2, (double *)(w + L.hp),
                                (double *)(w + L.hp2), (unsigned char *)(w + L.cp),
                                (unsigned *)(w + L.nnan), (const double *)(w + L.w), L.r, g,
                                g2, env, st),
           "envelope");
  }
  // initial state: sum |F0| over the core
  HIPCHK(hipMemcpyAsync(ft, ft0, N * 16, hipMemcpyDeviceToDevice, st), "copy");
  HIPCHK(odt::launch_apple_fourier(ft, ft0, nullptr, core, N, 1.0, 0,
                                   (double *)(w + L.part), (double *)(w + L.sum), st),
         "k_apple_fourier");
  double state = 0;
  HIPCHK(hipMemcpyAsync(&state, w + L.sum, 8, hipMemcpyDeviceToHost, st), "copy");
  HIPCHK(hipStreamSynchronize(st), "sync");
  double prev = state / (double)N, norm = 0;
  int ii = 0;
  for (; ii < max_iter; ++ii) {
    if (method == 0)
      HIPCHK(odt::launch_constraint_nn(data, n0, n1, n2, d_mask, bg_shell_width, st),
             "k_constraint_nn");
    else
      HIPCHK(odt::launch_constraint_sh(data, n0, n1, n2, d_mask, bg_shell_width,
                                       (double *)(w + L.shpart), (double *)(w + L.mm),
                                       (unsigned *)(w + L.hist), st),
             "k_sh");
    HIPCHK(hipMemcpyAsync(ft, data, N * 16, hipMemcpyDeviceToDevice, st), "copy");
    HIPCHK(fft3(ft, tmp, n0, n1, n2, -1, tw0, tw1, tw2, st), "fft3");
    HIPCHK(odt::launch_apple_fourier(ft, ft0, use_env ? env : nullptr, core, N,
                                     enforce_envelope, method == 1, (double *)(w + L.part),
                                     (double *)(w + L.sum), st),
           "k_apple_fourier");
    HIPCHK(hipMemcpyAsync(data, ft, N * 16, hipMemcpyDeviceToDevice, st), "copy");
    HIPCHK(fft3(data, tmp, n0, n1, n2, +1, tw0, tw1, tw2, st), "fft3");
    HIPCHK(odt::launch_scale_out(data, N, 1.0 / (double)N, 1, data, st), "scale");
    HIPCHK(hipMemcpyAsync(&state, w + L.sum, 8, hipMemcpyDeviceToHost, st), "copy");
    HIPCHK(hipStreamSynchronize(st), "sync");
    const double cur = state / (double)N;
    const double diff = cur - prev;
    if (ii == 0) {
      norm = diff;
    } else if (std::fabs(diff / norm) < min_diff) {
      break;
    }
    prev = cur;
  }
  if (iterations) *iterations = std::min(ii + 1, (int)max_iter);
  HIPCHK(odt::launch_scale_out(data, N, 1.0, out_precision == ODT_FP64, d_out, st), "out");
  return ODT_OK;
}

int odt_apple_envelope(int32_t n0, int32_t n1, int32_t n2, const void *d_ftdata,
                       const uint8_t *d_core, void *d_env, void *d_ws, size_t ws_bytes,
                       void *stream) {
  if (n0 < 1 || n1 < 1 || n2 < 1 || !d_ftdata || !d_core || !d_env || !d_ws)
    return fail(ODT_ERR_INVALID, "bad arguments");
  if (n0 != n2) return fail(ODT_ERR_INVALID, "`shape[0]` and `shape[2]` must be identical!");
  const AppleLayout L = apple_layout(n0, n1, n2);
  if (ws_bytes < L.total)
    return fail(ODT_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) +
                                       " bytes");
  hipStream_t st = (hipStream_t)stream;
  char *w = (char *)d_ws;
  const int mmax = std::max(n0, std::max(n1, n2));
  const long long N = (long long)n0 * n1 * n2;
  double *g = (double *)(w + L.tmp), *g2 = g + N;
  HIPCHK(odt::launch_gauss_weights((double *)(w + L.w), L.r, (double)mmax / 100.0, st),
         "k_gauss_weights");
  HIPCHK(hipMemsetAsync(w + L.nnan, 0, 4, st), "memset");
  HIPCHK(odt::launch_envelope((const cpx<double> *)d_ftdata, d_core, n0, n1, n2,
                              (double *)(w + L.hp), (double *)(w + L.hp2),
                              (unsigned char *)(w + L.cp), (unsigned *)(w + L.nnan),
                              (const double *)(w + L.w), L.r, g, g2, (double *)d_env, st),
         "envelope");
  return ODT_OK;
}

int odt_spline_prefilter(int precision, int order, int64_t nlines, int32_t n,
                         int64_t line_stride, int64_t elem_stride, void *d_data,
                         void *stream) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (order < 2 || order > 5) return fail(ODT_ERR_INVALID, "order must be 2..5");
  if (nlines < 0 || n < 1 || !d_data) return fail(ODT_ERR_INVALID, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipError_t e =
      precision == ODT_FP64
          ? odt::launch_iir<double>((double *)d_data, nlines, n, 1, 0, line_stride,
                                    elem_stride, order, st)
          : odt::launch_iir<float>((float *)d_data, nlines, n, 1, 0, line_stride,
                                   elem_stride, order, st);
  if (e != hipSuccess) return hip_fail(e, "k_iir");
  return ODT_OK;
}

int odt_rotate_accumulate(int precision, int order, int32_t nang, int32_t ny,
                          int32_t nz, int32_t nx, const void *d_coef,
                          const double *params, void *d_out, int accumulate,
                          void *stream) {
  if (precision != ODT_FP32 && precision != ODT_FP64)
    return fail(ODT_ERR_INVALID, "precision");
  if (order < 0 || order > 5) return fail(ODT_ERR_INVALID, "order must be 0..5");
  if (nang < 1 || ny < 1 || nz < 1 || nx < 1 || !d_coef || !params || !d_out)
    return fail(ODT_ERR_INVALID, "bad arguments");
  // the LDS windows of the tiles are sized for rotations: a row of the
  // matrix may not stretch a tile edge beyond 15*sqrt(2) input samples
  for (int a = 0; a < nang; ++a) {
    const double *m = params + 6 * (size_t)a;
    const double lim = std::sqrt(2.0) * (1 + 1e-12);
    if (!(std::fabs(m[0]) + std::fabs(m[1]) <= lim) ||
        !(std::fabs(m[2]) + std::fabs(m[3]) <= lim) || !std::isfinite(m[4]) ||
        !std::isfinite(m[5]))
      return fail(ODT_ERR_INVALID, "params must be rotation matrices");
  }
  hipStream_t st = (hipStream_t)stream;
  // parameters followed by the window-origin scratch of the tiles
  const size_t pbytes = align256(sizeof(double) * 6 * nang);
  const size_t wbytes = 4 * sizeof(int) * (size_t)nang * odt::rotate_window_entries(nz, nx);
  char *dp = nullptr;
  HIPCHK(dev_malloc((void **)&dp, pbytes + wbytes), "hipMalloc(params)");
  hipError_t e = hipMemcpyAsync(dp, params, sizeof(double) * 6 * nang,
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const int first = accumulate ? 0 : 1;
    const double *prm = (const double *)dp;
    int *win = (int *)(dp + pbytes);
    e = precision == ODT_FP64
            ? odt::launch_rotate_acc<double, double>(order, (const double *)d_coef, nang,
                                                     prm, win, (double *)d_out, nullptr,
                                                     ny, nz, nx, first, 1, st)
            : odt::launch_rotate_acc<float, float>(order, (const float *)d_coef, nang,
                                                   prm, win, (float *)d_out, nullptr, ny,
                                                   nz, nx, first, 1, st);
  }
  hipError_t e2 = hipStreamSynchronize(st);
  dev_free(dp);
  if (e != hipSuccess) return hip_fail(e, "k_rotate_acc");
  if (e2 != hipSuccess) return hip_fail(e2, "k_rotate_acc sync");
  return ODT_OK;
}

}  // extern "C"

// Internal interface between the plan (odt_plan.cpp) and the kernels
// (odt_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "odt_device.h"

namespace odt {

// Sizes shared by the transform kernels of one plan.
struct Geom {
  int A, Ny, Nx, ln;     // projections, detector rows/cols, z slices (= Nx)
  int lNy, lNx;          // padded transform sizes
  int logNy, logNx;      // log2 of the sizes, -1 if not a power of two
  int padyl, padxl;      // leading pad widths
  int pad_mode;          // 0 none, 1 edge, 2 linear ramp
  int Lx, Ly;            // support box (x- and y-frequencies kept)
  int jx0, jy0;          // signed frequency index of support column/row 0
  int fuse_x, fuse_y, fuse_z;  // prefilter folded into the transforms
  int col_mode;          // column stage: 0 default, 1 per angle, 2 angle loop
  double b1;             // B-spline sample at +-1 (1/8 order 2, 1/6 order 3)
  double z0;             // z of slice 0
};

// Per support point: fused z-prefilter terms (see k_col_ifft).
template <typename R>
struct PropAux {
  cpx<R> r0, r1;
  R ib;
  R pad;
};

template <typename R>
hipError_t launch_twiddles(cpx<R> *tw, int n, hipStream_t st);
// Twiddle table of a transform of length n, dft_table_len(n) entries (the
// Bluestein tables for lengths that are not powers of two; float32 tables
// are built in scratch64, dft_table_len(n) float64 entries, and rounded).
template <typename R>
hipError_t launch_dft_table(cpx<R> *tw, int n, cpx<double> *scratch64, hipStream_t st);
// Forward stage: pad, transforms and prefactor computed in C (float64 for
// both precisions where the transform fits the LDS in float64), the support
// spectrum S stored in R.
template <typename RI, typename C>
hipError_t launch_fwd_rows(const cpx<RI> *usin, const double *w, cpx<C> *F1,
                           const cpx<C> *twx, const Geom &g, double pr,
                           double pi, int a0, int na, int nb, hipStream_t st);
template <typename RI, typename C>
hipError_t launch_fwd_cols(const cpx<C> *F1, cpx<RI> *S, const cpx<C> *twy,
                           const cpx<C> *P, const Geom &g, int na, int nb,
                           hipStream_t st);
template <typename R>
hipError_t launch_col_ifft(const cpx<R> *S, cpx<R> *T, const cpx<R> *twy,
                           const double *pa, const PropAux<R> *aux,
                           const R *q0z, const R *q0y, const Geom &g, int na,
                           int nb, int ldt, const cpx<R> *gtab, hipStream_t st);
// Propagator table (ln x Lx x Ly) for the column stage.
template <typename R>
hipError_t launch_prop_table(const double *pa, const PropAux<R> *aux, const R *q0z,
                             const Geom &g, cpx<R> *out, hipStream_t st);
template <typename R, typename V>
hipError_t launch_row_ifft(const cpx<R> *T, V *slab, const cpx<R> *twx,
                           const R *q0x, const Geom &g, int na, int nb,
                           int slab0, int ldt, hipStream_t st,
                           const cpx<R> *S2 = nullptr, const cpx<R> *gt = nullptr);
// 2D plans (one detector row, propagator table present) whose row stage
// forms the column stage's product S * g_k on load (launch_row_ifft with S2
// and gt; no T, no column stage).
template <typename R>
bool row_fuses_2d(const Geom &g, bool gtab);
// Rotate-accumulate of nb angles; win is scratch of
// 4 * nb * rotate_window_entries(nz, nx) ints (window origin and extents per tile).
// groups > 1 (single-plane volumes): the angles run in that many concurrent
// groups, each summed into part (room for 2 * groups volumes of ny*nz*nx),
// then added to out in group order.
size_t rotate_window_entries(int nz, int nx);
template <typename R, typename V>
hipError_t launch_rotate_acc(int order, const V *coef, int nb,
                             const double *params, int *win, V *out, V *comp,
                             int ny, int nz, int nx, int first, int last,
                             hipStream_t st, int groups = 1, V *part = nullptr);
// Affine gather-accumulate of nb angles (tilted axis); win is scratch of
// 4 * nb * affine_window_entries(ny, nz, nx) ints (null: the direct gather).
size_t affine_window_entries(int ny, int nz, int nx);
template <typename R, typename V>
hipError_t launch_affine_acc(int order, const V *coef, int nb,
                             const double *params, int *win, V *out, V *comp, int ny,
                             int nz, int nx, int first, int last,
                             hipStream_t st);
template <typename R, typename V>
hipError_t launch_seq_rotate_sum(int order, V *slab, int nb,
                                 const double *params, V *out, V *comp,
                                 int ny, int nz, int nx, int first, int last,
                                 hipStream_t st);
// Poles and gain of the mirror-boundary B-spline prefilter of an order
// (2..5); false for other orders.
bool iir_poles(int order, int *np, double *z1, double *z2, double *gain);
template <typename R>
hipError_t launch_iir(R *data, long long nlines, int n, long long n0,
                      long long s0, long long s1, long long es, int order,
                      hipStream_t st);
template <typename R>
hipError_t launch_fft_batch(const cpx<R> *in, cpx<R> *out, long long batch,
                            int n, int logn, int dir, const cpx<R> *tw, int nb,
                            hipStream_t st);

// Object function to refractive index (odt_translate.hip); n elements
// (complex elements when is_complex), in place allowed.
template <typename R>
hipError_t launch_odt_to_ri(int is_complex, long long n, double km2, double nm,
                            const void *f, void *ri, hipStream_t st);
template <typename R>
hipError_t launch_opt_to_ri(int is_complex, long long n, double res, double nm,
                            const void *f, void *ri, hipStream_t st);

// Sinogram preparation (odt_prepare.hip): ndim 2 -> (A, W) with H = 1,
// ndim 3 -> (A, H, W); u0 null for u0 = 1, else broadcast to the full
// shape; rytov: complex output u0 (log|.| + i phase), else the real phase.
template <typename R>
hipError_t prepare_workspace(int ndim, int A, int H, int W, int rytov, size_t *bytes);
template <typename R>
hipError_t launch_prepare(int ndim, int A, int H, int W, const void *u, const void *u0,
                          int rytov, int align, void *out, void *ws, size_t ws_bytes,
                          hipStream_t st);

// Missing-apple-core correction (odt_apple.hip), complex128 working arrays.
hipError_t launch_apple_core(unsigned char *core, int n0, int n1, int n2, double km,
                             hipStream_t st);
hipError_t launch_apple_init(const void *f, int f64, cpx<double> *data, long long N,
                             hipStream_t st);
hipError_t launch_transpose3(const cpx<double> *in, cpx<double> *out, long long batch,
                             int rows, int cols, hipStream_t st);
hipError_t launch_envelope(const cpx<double> *F, const unsigned char *core, int n0, int n1,
                           int n2, double *hp, double *hp2, unsigned char *cp,
                           unsigned *nnan_dev, const double *w, int r, double *g, double *g2,
                           double *env, hipStream_t st);
hipError_t launch_constraint_nn(cpx<double> *data, int n0, int n1, int n2,
                                const unsigned char *mask, double bg_width, hipStream_t st);
// symmetric-histogram constraint; part: sh_part_bytes(), mm: 6 doubles,
// hist: sh_hist_bytes()
hipError_t launch_constraint_sh(cpx<double> *data, int n0, int n1, int n2,
                                const unsigned char *mask, double bg_width, double *part,
                                double *mm, unsigned *hist, hipStream_t st);
size_t sh_part_bytes();
size_t sh_hist_bytes();
// scipy.ndimage Gaussian taps (2 r + 1 doubles), normalised, on the device
hipError_t launch_gauss_weights(double *w, int r, double sigma, hipStream_t st);
hipError_t launch_apple_fourier(cpx<double> *F, cpx<double> *F0, const double *env,
                                const unsigned char *core, long long N, double factor,
                                int sh, double *part, double *sum, hipStream_t st);
hipError_t launch_scale_out(const cpx<double> *in, long long N, double s, int out64,
                            void *out, hipStream_t st);

// Adjoint stages (odt_adjoint.hip).
template <typename R>
hipError_t launch_adj_fir(const cpx<R> *in, cpx<R> *out, long long nlines, int n,
                          long long n0, long long s0, long long s1, long long es,
                          const R *hz, int K, hipStream_t st);
template <typename R>
hipError_t launch_adj_fir_rows(const cpx<R> *in, cpx<R> *out, long long nrows, int n,
                               const R *hz, int K, hipStream_t st);
// Zero padding on each side of the tap table of launch_adj_fir.
int adj_fir_pad();
template <typename R>
hipError_t launch_adj_prefilter_rows(cpx<R> *rows, long long nrows, int n,
                                     int order, hipStream_t st);
template <typename R>
hipError_t launch_adj_rotate(int order, const cpx<R> *g, int nb,
                             const double *params, cpx<R> *slab, int ny, int nz,
                             int nx, hipStream_t st);
// R_a^T from the planes-innermost gradient gt[z][x][y] (ny a multiple of 4).
template <typename R>
bool adj_rotate_t_supported(int order, int ny, int nz, int nx);
template <typename R>
hipError_t launch_adj_rotate_t(int order, const cpx<R> *gt, int nb, const double *params,
                               cpx<R> *slab, int ny, int nz, int nx, hipStream_t st);
// g[z][y][x] -> gt[z][x][y]
template <typename R>
hipError_t launch_adj_transpose(const cpx<R> *g, cpx<R> *gt, int ny, int nz, int nx,
                                hipStream_t st);
template <typename R>
hipError_t launch_scale_ends(R *data, long long nlines, int n, long long n0,
                             long long s0, long long s1, long long es, double s,
                             hipStream_t st);
template <typename R>
hipError_t launch_adj_fft(int stage, int n, const void *in, void *out,
                          const void *tw, const void *aux, const double *prop_a,
                          const double *weights, int a0, const Geom &g, int na,
                          int ldt, hipStream_t st);

}  // namespace odt

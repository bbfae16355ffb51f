/*
 * ves3d_hip.h -- C ABI of the MI355X (gfx950) implementation of the Ves3D
 * per-time-step hot path.
 *
 * Conventions
 *   - All pointers are device pointers unless a parameter says "host".
 *   - Sizes are element counts (not bytes) unless the name ends in _bytes.
 *   - Every entry point returns 0 on success or a code from the reference's
 *     Error_t numbering (include/Error.h:46-67 of lu1and10/ves3d).
 *   - `stream` is a hipStream_t passed as an opaque pointer (NULL = default
 *     stream).  Work is enqueued on that stream and the call returns without
 *     synchronising unless it says otherwise.
 *   - _f64 entries compute in FP64.  _f32 entries take FP32 buffers and compute
 *     in FP32 (the reference's single-precision build).  _f32c entries take
 *     FP64 buffers, compute the pair interactions in FP32 with compensated
 *     (TwoSum) accumulation of per-tile partial sums, and write FP64 results.
 *
 * Layouts follow the reference:
 *   AxisMajor vesicle fields  [ves][xyz][stride]       (Device<> kernels)
 *   point-major arrays        [point][xyz]             (StokesAlltoAll, PVFMMEval)
 */
#ifndef VES3D_HIP_H
#define VES3D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ves3d_stream_t;

/* Error_t values (include/Error.h:46-67). */
enum {
  VES3D_SUCCESS = 0,
  VES3D_UNKNOWN_ERROR = 1,
  VES3D_ARITHMETIC_ERROR = 2,
  VES3D_MEMORY_ERROR = 5,
  VES3D_NOT_IMPLEMENTED = 7,
  VES3D_INVALID_PARAMETER = 10,
  VES3D_INVALID_DEVICE = 20,
  VES3D_ACCURACY_ERROR = 30,
  VES3D_DIVERGENCE_ERROR = 31,
  VES3D_IO_ERROR = 50,
  VES3D_IO_BAD_STREAM = 51
};

/* ---- device and memory (Device<DT> ctor, Malloc/Free/Memcpy/Memset:
 *      include/Device.h:95-119) ------------------------------------------ */

/* Selects the device and checks that it is a gfx950 part.  Replaces the
 * Device<GPU> constructor (src/DeviceGPU.cc:1-21). */
int ves3d_hip_init(int device_id);

/* Writes the device name, CU count and clock (MHz) into host buffers. */
int ves3d_hip_device_info(char *name_host, size_t name_len, int *num_cu_host,
                          int *clock_mhz_host);

int ves3d_hip_malloc(void **ptr_host, size_t nbytes);
int ves3d_hip_free(void *ptr);
/* kind: 0 host->host, 1 host->device, 2 device->host, 3 device->device
 * (Device::MemcpyKind, include/Device.h:87-88).  Synchronous like the
 * reference's Device<GPU>::Memcpy. */
int ves3d_hip_memcpy(void *dst, const void *src, size_t nbytes, int kind);
int ves3d_hip_memset(void *ptr, int value, size_t nbytes);
int ves3d_hip_synchronize(void);

/* ---- Device<> elementwise / reduction / transpose / GEMM members
 *      (include/Device.h:121-232; semantics of src/DeviceCPU.cc:64-640).
 * Vector fields are AxisMajor [nv][3][stride], scalar fields [nv][stride].
 * Work is enqueued on the default stream; the scalar-returning calls
 * (maxabs, algebraic_dot, isnumeric) wait for their result.  NULL y / w
 * arguments select the reference's "x only" branches. */
int ves3d_hip_dot_product_f64(const double *u, const double *v, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_dot_product_f32(const float *u, const float *v, size_t stride, size_t n_vecs, float *out);
int ves3d_hip_cross_product_f64(const double *u, const double *v, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_cross_product_f32(const float *u, const float *v, size_t stride, size_t n_vecs, float *out);
int ves3d_hip_sqrt_f64(const double *x, size_t length, double *out);
int ves3d_hip_sqrt_f32(const float *x, size_t length, float *out);
/* ax: out[n][i] = a[i] x[n][i] (a has `stride` entries) */
int ves3d_hip_ax_f64(const double *a, const double *x, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_ax_f32(const float *a, const float *x, size_t stride, size_t n_vecs, float *out);
int ves3d_hip_xy_f64(const double *x, const double *y, size_t length, double *out);
int ves3d_hip_xy_f32(const float *x, const float *y, size_t length, float *out);
/* xyInv: out = x / y (x == NULL: 1 / y) */
int ves3d_hip_xyinv_f64(const double *x, const double *y, size_t length, double *out);
int ves3d_hip_xyinv_f32(const float *x, const float *y, size_t length, float *out);
/* uyInv: vector u divided by scalar y per point */
int ves3d_hip_uyinv_f64(const double *u, const double *y, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_uyinv_f32(const float *u, const float *y, size_t stride, size_t n_vecs, float *out);
/* axpy: out = a x + y (y == NULL: a x) */
int ves3d_hip_axpy_f64(double a, const double *x, const double *y, size_t length, double *out);
int ves3d_hip_axpy_f32(float a, const float *x, const float *y, size_t length, float *out);
/* apx: out[n][i] = a[n] + x[n][i] */
int ves3d_hip_apx_f64(const double *a, const double *x, size_t stride, size_t n_subs, double *out);
int ves3d_hip_apx_f32(const float *a, const float *x, size_t stride, size_t n_subs, float *out);
/* avpw: out = a[vec] v + w (vector fields, w may be NULL) */
int ves3d_hip_avpw_f64(const double *a, const double *v, const double *w, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_avpw_f32(const float *a, const float *v, const float *w, size_t stride, size_t n_vecs, float *out);
/* xvpw: out = x (scalar field) v + w (vector fields, w may be NULL) */
int ves3d_hip_xvpw_f64(const double *x, const double *v, const double *w, size_t stride, size_t n_vecs, double *out);
int ves3d_hip_xvpw_f32(const float *x, const float *v, const float *w, size_t stride, size_t n_vecs, float *out);
/* Reduce: out[ii][kk] = sum_j x[ii][kk][j] w[ii][j] qw[j] (x == NULL: sum w qw) */
int ves3d_hip_reduce_f64(const double *x, int x_dim, const double *w, const double *quad_w, size_t stride, size_t ns, double *out);
int ves3d_hip_reduce_f32(const float *x, int x_dim, const float *w, const float *quad_w, size_t stride, size_t ns, float *out);
int ves3d_hip_maxabs_f64(const double *x, size_t length, double *result_host);
int ves3d_hip_maxabs_f32(const float *x, size_t length, float *result_host);
/* AlgebraicDot with FP64 accumulation (the reference's GPU form only reduces
 * block 0, src/CudaKernels.cu:1125) */
int ves3d_hip_algebraic_dot_f64(const double *x, const double *y, size_t length, double *result_host);
int ves3d_hip_algebraic_dot_f32(const float *x, const float *y, size_t length, float *result_host);
/* isNumeric: 1 if every entry is finite.  (The reference's CPU version
 * starts from false and ANDs, so it can never fail, src/DeviceCPU.cc:632-638.) */
int ves3d_hip_isnumeric_f64(const double *x, size_t length, int *result_host);
int ves3d_hip_isnumeric_f32(const float *x, size_t length, int *result_host);
/* Transpose: out[w][h] = in[h][w] */
int ves3d_hip_transpose_f64(const double *in, size_t height, size_t width, double *out);
int ves3d_hip_transpose_f32(const float *in, size_t height, size_t width, float *out);
/* gemm: BLAS column-major C = alpha op(A) op(B) + beta C ('N' / 'T'). */
int ves3d_hip_gemm_f64(char transA, char transB, int m, int n, int k, double alpha,
                       const double *A, int lda, const double *B, int ldb, double beta,
                       double *C, int ldc);
int ves3d_hip_gemm_f32(char transA, char transB, int m, int n, int k, float alpha,
                       const float *A, int lda, const float *B, int ldb, float beta,
                       float *C, int ldc);

/* Selects the kernel behind every grouped GEMM (gemm above, the SH
 * transforms, rotations, singular-operator build and the learned
 * preconditioner's network): 0 = VALU FMA tiles, 1 = MFMA
 * (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, the default).  No
 * counterpart in the reference (its GEMMs are BLAS / PVFMM calls); the
 * environment variable VES3D_GEMM=valu sets the initial value to 0. */
int ves3d_hip_set_gemm_kernel(int kind);

/* Output tile of the MFMA grouped GEMM: -1 (the default) picks per
 * descriptor list the shape that minimises padded work plus operand
 * traffic; 0..8 force one of the shapes (rows x columns) 64x64, 64x16,
 * 128x16, 256x16, 64x32, 128x32, 64x48, 128x48, 128x64.  For tests and
 * tuning; no counterpart in the reference. */
int ves3d_hip_set_gemm_tile(int tile);

/* The sustained VALU rate of this GPU: lane fused multiply-adds per second
 * of independent FMA chains, FP32 non-packed (fp64 = 0) or FP64 (fp64 = 1),
 * the reference for the far-field kernels' rooflines.  Synchronous.  No
 * counterpart in the reference. */
int ves3d_hip_valu_probe(int fp64, double *lane_fma_per_s);

/* Selects the kernel of the far field between vesicles inside the implicit
 * step (StokesVelocity's PVFMMEval call, src/StokesVelocity.cc:481-485):
 * 0 = sources staged through LDS tiles, 1 = sources read through the scalar
 * cache (the default), 2 = as 1 with two targets per lane in the FP32 form.
 * All give the same sums up to rounding order.  No counterpart in the
 * reference. */
int ves3d_hip_set_far_kernel(int kind);

/* ---- direct Stokes sums ------------------------------------------------ */

/* Device<DT>::DirectStokes (include/Device.h:205-209).  For each vesicle vt
 * and target index t in [trg_idx_head, trg_idx_tail):
 *   pot = 1/(8 pi) sum_s [ (w_s d_s)/r + ((w_s d_s).r) r / r^3 ],  r = src-trg,
 * over the sources of the same vesicle; qw == NULL means w_s = 1
 * (src/DeviceCPU.cc:513-541).  src,den: [n_surfs][3][src_stride];
 * trg,pot: [n_surfs][3][trg_stride]; qw: [src_stride]. Coincident points
 * (r = 0) contribute nothing (src/CPUKernels.cc:37). */
int ves3d_hip_direct_stokes_f64(const double *src, const double *den,
                                const double *qw, size_t src_stride,
                                size_t trg_stride, size_t n_surfs,
                                const double *trg, size_t trg_idx_head,
                                size_t trg_idx_tail, double *pot,
                                ves3d_stream_t stream);
int ves3d_hip_direct_stokes_f32(const float *src, const float *den,
                                const float *qw, size_t src_stride,
                                size_t trg_stride, size_t n_surfs,
                                const float *trg, size_t trg_idx_head,
                                size_t trg_idx_tail, float *pot,
                                ves3d_stream_t stream);

/* Device<DT>::DirectStokesDoubleLayer (include/Device.h:211-215):
 *   pot = -3/(4 pi) sum_s w_s (r.n_s)(r.d_s) r / r^5,  r = src - trg
 * (src/CPUKernels.cc:1059-1136).  norm: [n_surfs][3][src_stride]. */
int ves3d_hip_direct_stokes_dl_f64(const double *src, const double *norm,
                                   const double *den, const double *qw,
                                   size_t src_stride, size_t trg_stride,
                                   size_t n_surfs, const double *trg,
                                   size_t trg_idx_head, size_t trg_idx_tail,
                                   double *pot, ves3d_stream_t stream);
int ves3d_hip_direct_stokes_dl_f32(const float *src, const float *norm,
                                   const float *den, const float *qw,
                                   size_t src_stride, size_t trg_stride,
                                   size_t n_surfs, const float *trg,
                                   size_t trg_idx_head, size_t trg_idx_tail,
                                   float *pot, ves3d_stream_t stream);

/* StokesAlltoAll (include/CPUKernels.h:46-49, src/CPUKernels.cc:966-1054),
 * the VesInteraction<T>::InteractionFun_t without PVFMM
 * (include/VesInteraction.h:21): single-layer sum over all np point-major
 * sources/targets, coincident pairs skipped. src, den, pot: [np][3]. */
int ves3d_hip_stokes_all2all_f64(const double *src, const double *den,
                                 size_t np, double *pot, ves3d_stream_t stream);
int ves3d_hip_stokes_all2all_f32c(const double *src, const double *den,
                                  size_t np, double *pot,
                                  ves3d_stream_t stream);

/* Direct-sum stand-in for PVFMMEval(src_pos, sl_den, dl_den, n_src, trg_pos,
 * trg_vel, n_trg, ctx, setup) (include/PVFMMInterface.h:53), with the pair
 * kernels of src/PVFMMInterface.cc:161-238 and :284-369:
 *   trg_vel = (1/8pi) sum [f/r + (f.d)d/r^3] + (3/4pi) sum (d.n)(d.g)d/r^5,
 * d = trg - src, pairs with |d|^2 <= 128*eps(double) skipped.
 * src_pos: [n_src][3]; sl_den: [n_src][3] or NULL; dl_den: [n_src][6] laid
 * out [g (3), n (3)] per source (qforce_double, src/StokesVelocity.cc:351-372)
 * or NULL; trg_pos, trg_vel: [n_trg][3] (trg_vel is overwritten).
 * Grouping (an extension for the vesicle far field): when src_per_group and
 * trg_per_group are both > 0, sources and targets are in n_src/src_per_group
 * consecutive groups (vesicles) and pairs inside one group are skipped --
 * the own-vesicle part that src/NearSingular.cc:1143-1212 subtracts again. */
int ves3d_hip_direct_eval_f64(const double *src_pos, const double *sl_den,
                              const double *dl_den, size_t n_src,
                              const double *trg_pos, double *trg_vel,
                              size_t n_trg, size_t src_per_group,
                              size_t trg_per_group, ves3d_stream_t stream);
int ves3d_hip_direct_eval_f32c(const double *src_pos, const double *sl_den,
                               const double *dl_den, size_t n_src,
                               const double *trg_pos, double *trg_vel,
                               size_t n_trg, size_t src_per_group,
                               size_t trg_per_group, ves3d_stream_t stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ves3d_hip_direct_eval_f64 with the same arguments and
 * results, whose one scratch buffer (the per-slice FP64 partial sums) is a
 * fixed 64 MiB device array per GPU, held until the process ends, so the
 * call performs no hipMalloc / hipFree.  n_trg may be at most 2,796,202
 * (n_trg * 24 bytes <= 64 MiB), else VES3D_INVALID_PARAMETER; where the
 * partials of the default slice count do not fit, fewer slices are used
 * (the same sums in another order).  All calls on one GPU share the array:
 * calls that overlap on that GPU (different streams, threads) race. */
int ves3d_hip_direct_eval_f64_noalloc(const double *src_pos, const double *sl_den,
                                      const double *dl_den, size_t n_src,
                                      const double *trg_pos, double *trg_vel,
                                      size_t n_trg, size_t src_per_group,
                                      size_t trg_per_group, ves3d_stream_t stream);

/* ---- spherical-harmonic transforms (SphericalHarmonics<Real>,
 *      include/SphericalHarmonics.h:14-24, src/SphericalHarmonics.cc:3-360) --
 * Grid layout: per function (p+1) x 2p values, latitude i = i-th Gauss-Legendre
 * node (ascending), longitude j = pi j / p.  Coefficient layout: per function
 * p(p+2) values in blocks j = 0..2p-1 of length p+1-(j+1)/2 (order (j+1)/2,
 * cosine for odd j and j=0, sine for even j>0, degree ascending).  N is the
 * number of functions.  The f32 forms run the same transforms in FP32. */

/* SHC2Grid: S [N][p0(p0+2)] -> X [N][(p1+1) 2p1].  X_theta (longitude
 * derivative, MatFourierGrad) and X_phi (latitude derivative, MatLegendreGrad)
 * are optional outputs (NULL to skip), named as in the reference. */
int ves3d_hip_shc2grid_f64(const double *S, size_t N, int p0, int p1, double *X,
                           double *X_theta, double *X_phi, ves3d_stream_t stream);
int ves3d_hip_shc2grid_f32(const float *S, size_t N, int p0, int p1, float *X,
                           float *X_theta, float *X_phi, ves3d_stream_t stream);
/* Grid2SHC: X [N][(p0+1) 2p0] -> S [N][p1(p1+2)]. */
int ves3d_hip_grid2shc_f64(const double *X, size_t N, int p0, int p1, double *S,
                           ves3d_stream_t stream);
int ves3d_hip_grid2shc_f32(const float *X, size_t N, int p0, int p1, float *S,
                           ves3d_stream_t stream);
/* SHC2GridTranspose (the adjoint of SHC2Grid(p1 -> p0)):
 * X [N][(p0+1) 2p0] -> S [N][p1(p1+2)]. */
int ves3d_hip_shc2grid_transpose_f64(const double *X, size_t N, int p0, int p1,
                                     double *S, ves3d_stream_t stream);
int ves3d_hip_shc2grid_transpose_f32(const float *X, size_t N, int p0, int p1,
                                     float *S, ves3d_stream_t stream);
/* SHC2Pole: S [N][p0(p0+2)] -> P [N][2], values at x = -1 and x = +1. */
int ves3d_hip_shc2pole_f64(const double *S, size_t N, int p0, double *P,
                           ves3d_stream_t stream);
int ves3d_hip_shc2pole_f32(const float *S, size_t N, int p0, float *P,
                           ves3d_stream_t stream);

/* ---- device SHT in the SHTrans layout (src/SHTrans.cc) -------------------
 * Coefficients are "order-major": coefficient block j (order (j+1)/2) of all
 * nfun functions in turn, degrees ascending within a block -- the layout of
 * the reference's Galerkin vectors (src/SHTrans.cc:58-120).  p(p+2)
 * coefficients per function.  FP64. */

/* SHTrans::forward: grid [nfun][(p+1) 2p] -> shc (order-major). */
int ves3d_hip_shtrans_forward(int p, size_t nfun, const double *grid, double *shc,
                              ves3d_stream_t stream);
/* SHTrans::backward (du = dv = 0), backward_du (1,0), backward_dv (0,1),
 * backward_d2u (2,0), backward_d2v (0,2), backward_duv (1,1): u is the
 * latitude direction (d1legTrans / d2legTrans matrices), v the longitude. */
int ves3d_hip_shtrans_backward(int p, size_t nfun, const double *shc, int du, int dv,
                               double *grid, ves3d_stream_t stream);
/* SHTrans::ScaleFreq: shc_out = scale * shc_in, scale [p(p+2)] given per
 * coefficient of one function (function-major order, as the reference's
 * filter and preconditioner arrays). */
int ves3d_hip_shtrans_scale_freq(int p, size_t nfun, const double *shc_in,
                                 const double *scale, double *shc_out,
                                 ves3d_stream_t stream);
/* SHTrans::lowPassFilter with the box filter of frequency freq (degrees
 * n <= freq kept, src/SHTrans.cc:23-31): grid -> grid. */
int ves3d_hip_shtrans_low_pass(int p, size_t nfun, int freq, const double *in,
                               double *out, ves3d_stream_t stream);
/* Resample (src/HelperFuns.cc:436-493): grid at order p -> grid at order q. */
int ves3d_hip_resample(int p, int q, size_t nfun, const double *in, double *out,
                       ves3d_stream_t stream);

/* ---- rotations and the singular self-interaction ------------------------ */

/* RotateAll (src/SphericalHarmonics.cc:362-449): S [N*dof][p(p+2)] ->
 * S_out [N][p+1][p][dof][p(p+2)], the coefficients rotated so that grid point
 * (t, j) sits at the pole. */
int ves3d_hip_rotate_all_f64(const double *S, size_t N, int p, int dof,
                             double *S_out, ves3d_stream_t stream);
/* RotateTranspose (src/SphericalHarmonics.cc:451-541): the adjoint, same
 * [N][p+1][p][dof][p(p+2)] layout in and out. */
int ves3d_hip_rotate_transpose_f64(const double *S_in, size_t N, int p, int dof,
                                   double *S_out, ves3d_stream_t stream);

/* Kernel choice of the singular-operator build (no counterpart in the
 * reference, global per process): 1 (the default) evaluates the FP64 build's
 * pole kernel and its transposed synthesis (Fourier stage on FP64 MFMA,
 * Legendre stage) in one kernel, where p1 = 2 p0 and p0 <= 16; 0 runs them
 * as the pole kernel and two grouped GEMMs with the kernel values in HBM
 * between them (the FP32 build always does). */
int ves3d_hip_set_pole_kernel(int fused);
/* The fused stage's launch on the current device for orders (p0, p1) of the
 * FP64 build (no counterpart in the reference): latitude rows per pass,
 * threads per workgroup and the persistent grid's workgroups; all 0 where
 * the build runs the separate kernels. */
int ves3d_hip_pole_kernel_shape(int p0, int p1, int *rows_host, int *threads_host,
                                int *workgroups_host);
/* StokesSingularInteg (include/SphericalHarmonics.h:40,
 * src/SphericalHarmonics.cc:862-886, :952-1180): per-vesicle single- and
 * double-layer self-interaction operators on SH coefficients.
 * X0: vesicle positions [nves][3][(p0+1) 2p0] (FP64); p1 the upsampled order
 * of the quadrature (the reference uses 2 p0).  SL, DL: [nves][3 Nc][3 Nc],
 * Nc = p0(p0+2), row = density coefficient (component-major), column =
 * velocity coefficient; either may be NULL.  The f32 form runs the transform
 * GEMMs in FP32 and the pole-kernel evaluation in FP64, and stores FP32. */
int ves3d_hip_singular_build_f64(const double *X0, size_t nves, int p0, int p1,
                                 double *SL, double *DL, ves3d_stream_t stream);
int ves3d_hip_singular_build_f32(const double *X0, size_t nves, int p0, int p1,
                                 float *SL, float *DL, ves3d_stream_t stream);

/* Self-interaction apply (src/StokesVelocity.cc:397-447):
 * V[v] = F[v] SL[v] + G[v] DL[v] (row vector times matrix) for every vesicle;
 * F, G, V: [nves][3 Nc] FP64 coefficient vectors; SL or DL (and its vector)
 * may be NULL.  The f32 form reads FP32 matrices, accumulates in FP64. */
int ves3d_hip_self_apply_f64(const double *SL, const double *F, const double *DL,
                             const double *G, size_t nves, int p, double *V,
                             ves3d_stream_t stream);
int ves3d_hip_self_apply_f32(const float *SL, const double *F, const float *DL,
                             const double *G, size_t nves, int p, double *V,
                             ves3d_stream_t stream);

/* ---- implicit solver and time step -------------------------------------- */

/* Linear operator / preconditioner callback on device vectors, the device
 * counterpart of ParallelLinOp<T>::apply_type and
 * ParallelLinSolver<T>::precond_type (include/ParallelLinSolverInterface.h:111,
 * :160): y = Op(x); returns 0 on success. */
typedef int (*ves3d_apply_fn)(const void *ctx, const double *x, double *y);

/* BiCGStab<Container, MatVec, Precond>::operator() (include/BiCGStab.h:62-65,
 * src/BiCGStab.cc:45-138): right-preconditioned BiCGStab on device vectors of
 * length n with FP64 inner products.  P may be NULL (identity).  On return
 * *max_iter holds the iterations used, *tol the achieved relative residual
 * and *bicgs_status a BiCGSReturn code (100 success, 101 max iterations,
 * 102 rho breakdown, 103 omega breakdown, 104 NaN residual). */
int ves3d_hip_bicgstab(size_t n, ves3d_apply_fn A, const void *A_ctx,
                       ves3d_apply_fn P, const void *P_ctx, double *x,
                       const double *b, int *max_iter, double *tol,
                       int *bicgs_status, ves3d_stream_t stream);
/* ParallelLinSolver<T>::Solve of the PETSc implementation (KSP GMRES with a
 * shell preconditioner, src/ParallelLinSolver_Petsc.cc:545-551, :600-634) as
 * flexible GMRES: right-preconditioned, restarted every `restart` (1..63)
 * iterations, the preconditioned vectors kept so that P may change between
 * applications.  Classical Gram-Schmidt applied twice with FP64 dots.  On
 * return *max_iter holds the iterations (operator applications after the
 * initial residual), *tol the relative residual ||b - A x|| / ||b|| formed
 * explicitly from the Arnoldi basis, *status 100 (converged), 101 (iteration
 * limit) or 104 (NaN).  Allocates (2 restart + 1) n doubles of scratch. */
int ves3d_hip_fgmres(size_t n, ves3d_apply_fn A, const void *A_ctx, ves3d_apply_fn P,
                     const void *P_ctx, double *x, const double *b, int *max_iter, double *tol,
                     int *status, int restart, ves3d_stream_t stream);

/* ---- StokesVelocity<Real> (src/StokesVelocity.cc:146-539) --------------
 * The Stokes velocity of a single-layer density f and a double-layer
 * density d on nves vesicles at order p (SH grid points [nves][3][(p+1)2p],
 * AxisMajor), upsampled to p_up (0: 2p) for the far field and the
 * near-singular correction, in FP64: the singular self term (the SL and DL
 * self operators of StokesSingularInteg), the direct sum over the other
 * vesicles' upsampled sources that stands in for PVFMMEval, and the
 * near-singular correction (NearSingular; r_near <= 0: the reference's
 * r_ves 2/sqrt(p_up)) without repulsion.  create / destroy a handle
 * (the singular operators of nves vesicles: 2 x 72 p^2 (p+2)^2 bytes per
 * vesicle); set_src = SetSrcCoord (synchronous; builds the operators and
 * the near-pair setup); eval = SetDensitySL(f_sl) + SetDensityDL(f_dl) +
 * operator(): trg == NULL gives the velocity on the surfaces themselves
 * (trg_is_surf), vel [nves][3][(p+1)2p]; otherwise the velocity at n_trg
 * targets off the surfaces (SetTrgCoord), trg and vel point-major
 * [n_trg][3], without the self term.  Either density may be NULL, not
 * both.  eval is synchronous. */
int ves3d_hip_stokes_create(int p, int p_up, size_t nves, double r_near, void **handle_host);
int ves3d_hip_stokes_destroy(void *handle);
int ves3d_hip_stokes_set_src(void *handle, const double *X, ves3d_stream_t stream);
int ves3d_hip_stokes_eval(void *handle, const double *f_sl, const double *f_dl,
                          const double *trg, size_t n_trg, double *vel, ves3d_stream_t stream);

/* Parameters of the globally implicit interfacial-velocity step
 * (InterfacialVelocity::updateImplicit with solve_for_velocity,
 * src/InterfacialVelocity.cc:203-560): shear background flow
 * u = (shear_rate z, 0, 0); per-vesicle properties (viscosity contrast,
 * excess density under gravity, bending moduli) are set by
 * ves3d_hip_stepper_set_vesicle_props. */
typedef struct ves3d_step_params {
  size_t nves;            /* number of vesicles */
  int p;                  /* SH order of the surfaces */
  int p_up;               /* upsampled order (0: 2p, the reference default) */
  int filter_freq;        /* low-pass filter frequency (0: 2p/3) */
  double dt;              /* time step */
  double shear_rate;      /* background shear rate */
  double bending_modulus; /* kappa_b, the same for every vesicle */
  double tol;             /* BiCGStab relative residual tolerance */
  int max_iter;           /* BiCGStab iteration limit */
  int precision;          /* 0: FP64; 1: FP32 singular operator + FP32-compensated far field;
                             2: FP64 singular operator + FP32-compensated far field */
  int precond;            /* 0: none; 1: diagonal spectral (ConfigurePrecond); 2: learned */
  int near_off;           /* 1 disables the near-singular correction and the repulsion force */
  double repul_dist;      /* repulsion length (0: the reference default 5e-2; < 0: none) */
  double r_near;          /* near-zone radius (0: r_ves 2/sqrt(p_up), src/NearSingular.cc:83) */
  int post_off;           /* 1 skips the reparametrization and area / volume correction
                             that follow the position update (src/EvolveSurface.cc:330-335) */
  int inner_max_iter;     /* > 0: the preconditioner becomes an inner BiCGStab on the
                             vesicles' own (block-diagonal) operator, itself preconditioned
                             by `precond` (diagonal or learned), stopped at inner_tol or
                             inner_max_iter iterations.  0 (the reference's ImplicitPrecond
                             form): `precond` is applied once.  No counterpart in the
                             reference. */
  double inner_tol;       /* relative residual of the inner solve (0: 1e-3) */
  int extrapolate;        /* 1: the solve starts from 2 x_{n-1} - x_{n-2}, the linear
                             extrapolation of the last two steps' solutions; 0: from
                             x_{n-1}, as the reference's AssembleInitial
                             (src/InterfacialVelocity.cc:873-919).  No counterpart in
                             the reference. */
  double far_coarse;      /* > 1: two-level far field.  A wave of 64 targets reads a
                             vesicle's sources on its own p grid (with that grid's
                             quadrature weights) instead of its upsampled ones when
                             every target of the wave is at least far_coarse times the
                             vesicle's radius from its centre (the p-grid quadrature is
                             then as accurate as the upsampled one).  0: upsampled
                             sources for every pair, the reference's far field
                             (src/StokesVelocity.cc:477-490).  No counterpart in the
                             reference, whose FMM makes its own approximation there. */
  double far_coarse2;     /* > far_coarse (needs far_coarse): three-level far field.
                             Beyond far_coarse2 radii a wave reads the vesicle at order
                             p/2: the degree-p surface at the order-p/2 grid points with
                             the density times the area element (per unit-sphere
                             measure) truncated to degree p/2 and the order-p/2
                             Gauss-Legendre weights.  0: two levels.  No counterpart in
                             the reference. */
  int fallback;           /* when a learned solve fails or its true residual misses the
                             tolerance, the step solves again from the same initial
                             guess: 0 with the diagonal spectral preconditioner
                             (ImplicitPrecond), 1 unpreconditioned. */
  int solver;             /* outer solver: 0 right-preconditioned BiCGStab
                             (src/BiCGStab.cc:45-138) with a true-residual check;
                             1 flexible GMRES (ves3d_hip_fgmres), the GMRES of the
                             reference's globally implicit solve
                             (src/ParallelLinSolver_Petsc.cc:545-551) in the form a
                             varying preconditioner needs. */
  int restart;            /* FGMRES restart length (0: 30; at most 63) */
} ves3d_step_params;

typedef struct ves3d_step_stats {
  int iterations;         /* outer solver iterations (FGMRES: operator applications
                             after the initial residual) */
  int matvecs;            /* operator applications (incl. residual checks) */
  int status;             /* BiCGSReturn code */
  int fell_back;          /* 1 if the learned solve failed or missed the tolerance
                             and was repeated with the diagonal preconditioner */
  double relres;          /* recursive relative residual at exit */
  double true_relres;     /* ||b - A x|| / ||b|| of the returned solution */
  double ms_setup, ms_rhs, ms_solve, ms_total;
  double ms_far;          /* summed device time of the far-field direct sums */
  int far_calls;          /* number of far-field direct sums in the step */
  int near_pairs;         /* near-singular (target, vesicle) pairs of the step */
  double r_near;          /* near-zone radius of the step */
  int reparam_iters;      /* reparametrization iterations (0 with post_off) */
  int av_iters;           /* area / volume correction iterations */
  double ms_post;         /* wall time of the two */
  double ms_far2;         /* summed device time of the two-density far-field passes
                             (the right-hand side together with the first matvec) */
  int far2_calls;         /* number of two-density passes */
  int inner_iters;        /* inner BiCGStab iterations summed over the step's
                             preconditioner applications (0 without inner solves) */
  int local_matvecs;      /* local-operator applications of the inner solves */
  double far_pairs;       /* source-target pairs of one far-field evaluation of one
                             density in this step (fewer with far_coarse) */
  double ms_near;         /* summed device time of the near-singular corrections */
} ves3d_step_stats;

/* Creates / destroys a stepper (device state for nves vesicles). */
int ves3d_hip_stepper_create(const ves3d_step_params *params, void **handle_host);
int ves3d_hip_stepper_destroy(void *handle);
/* Length of the Galerkin solver vector, 4 nves p(p+2) (velocity block of
 * 3 nves functions then tension block of nves functions, each in the
 * order-major SHTrans layout, src/InterfacialVelocity.cc:562-573). */
int ves3d_hip_stepper_size(void *handle, size_t *n_host);
/* Copies the state in / out: positions X [nves][3][(p+1)2p], interfacial
 * velocity U (same layout) and tension sigma [nves][(p+1)2p]; any may be NULL. */
int ves3d_hip_stepper_set_state(void *handle, const double *X, const double *U,
                                const double *sigma);
int ves3d_hip_stepper_get_state(void *handle, double *X, double *U, double *sigma);
/* Geometry, singular operators and far-field setup for the current positions
 * (Prepare + StokesVelocity setup, src/InterfacialVelocity.cc:576-611). */
int ves3d_hip_stepper_prepare(void *handle);
/* AssembleRhsVel (src/InterfacialVelocity.cc:715-787) into b [n]. */
int ves3d_hip_stepper_rhs(void *handle, double *b);
/* ImplicitApply (src/InterfacialVelocity.cc:981-1052): y = A x. */
int ves3d_hip_stepper_matvec(void *handle, const double *x, double *y);
/* AssembleRhsVel into b and ImplicitApply y = A x together, the two Stokes
 * evaluations sharing one far-field pass (what each step does for its right-
 * hand side and the solver's first matvec). */
int ves3d_hip_stepper_rhs_matvec(void *handle, const double *x, double *b, double *y);
/* ImplicitPrecond with the diagonal spectral preconditioner
 * (src/InterfacialVelocity.cc:1054-1109): y = D x. */
int ves3d_hip_stepper_precond(void *handle, const double *x, double *y);
/* The inner-solve preconditioner of the step (ves3d_step_params.inner_max_iter
 * > 0): y ~ A_loc^-1 x by BiCGStab on every vesicle's own operator from
 * y = 0, preconditioned by the diagonal (learned = 0) or the attached learned
 * preconditioner (learned = 1), to inner_tol (1e-3 if 0) within
 * inner_max_iter (50 if 0) iterations; *iters_host gets the iteration
 * count.  Requires a prepared step.  No counterpart in the reference. */
int ves3d_hip_stepper_precond_inner(void *handle, const double *x, double *y, int learned,
                                    int *iters_host);
/* The relative residual |b - A x| / |b| of the current state: the step's
 * right-hand side b and operator A at the positions, x the packed velocity
 * and tension (the true-residual check of src/InterfacialVelocity.cc:181-188
 * for a given solution).  With a precision-0 stepper it judges a solution
 * of any form against the FP64 operator. */
int ves3d_hip_stepper_residual(void *handle, double *relres_host);
/* Source-target pairs of one far-field evaluation (one density) as of the
 * last prepare: every target against every other vesicle's upsampled
 * sources, or fewer with far_coarse.  No counterpart in the reference. */
int ves3d_hip_stepper_far_pairs(void *handle, double *pairs_host);
/* No counterpart in the reference (it feeds the learned preconditioner):
 * ImplicitApply with every vesicle's own operator only, the singular
 * self-interaction without the far field and the near-singular correction:
 * y = A_loc x. */
int ves3d_hip_stepper_matvec_local(void *handle, const double *x, double *y);
/* The attached, trained learned preconditioner (see ves3d_hip_lpc_create):
 * y = P(x); the stepper must be prepared. */
int ves3d_hip_stepper_precond_learned(void *handle, const double *x, double *y);
/* The vesicles' canonical frames of the last prepare into frames [nv][3][3]
 * (device, row-major): the orthogonal polar factor of
 * C = sum_k w_k X(k) s(k)^T over the grid, s the parameter sphere
 * (sin t cos p, sin t sin p, cos t) and w the Gauss-Legendre weights. */
int ves3d_hip_stepper_frames(void *handle, double *frames);
/* One full implicit step: prepare, right-hand side, BiCGStab from the
 * previous (velocity, tension) as initial guess, true-residual check,
 * X += dt U, then (unless post_off) InterfacialVelocity::reparam and
 * EvolveSurface::AreaVolumeCorrection against the area and volume recorded
 * by the last set_state with positions.  stats_host may be NULL.
 * A solution is accepted when BiCGStab converged and its true residual is
 * at most 1.05 tol; BiCGStab continues from its own result (at most twice,
 * within the iteration budget) when only the true residual misses.  A
 * learned solve that is not accepted is repeated from the same initial
 * guess with the diagonal spectral preconditioner.  When no solution is
 * accepted the step returns VES3D_DIVERGENCE_ERROR
 * (src/InterfacialVelocity.cc:176-188), fills the statistics and leaves the
 * positions, velocity and tension as they were. */
int ves3d_hip_stepper_step(void *handle, ves3d_step_stats *stats_host);
/* InterfacialVelocity::reparam without upsampling (src/InterfacialVelocity.cc:
 * 1639-1773) on the current positions, with the reference's default
 * parameters (PolyKReparam, exponent 4, rep_ts 1e-3, rep_tol 1e-4, at most 10
 * iterations, src/Parameters.cc:39-45).  Synchronous. */
int ves3d_hip_stepper_reparam(void *handle, int *iters_host);
/* EvolveSurface::AreaVolumeCorrection (src/EvolveSurface.cc:352-470): Newton
 * steps on the upsampled surface along (-2H n, n) until area and volume
 * match the recorded ones to 1e-12 (at most 10).  Synchronous. */
int ves3d_hip_stepper_area_volume(void *handle, int *iters_host);
/* Per-vesicle properties (VesicleProperties, src/VesicleProps.cc:19-55, as
 * read by ves3d_read_vesicle_props): bending moduli, viscosity contrasts and
 * excess densities, host arrays [nves], and the gravity field (host [3],
 * Parameters::gravity_field, default (0, 0, -1)); NULL leaves an entry as it
 * is.  The step's forces then carry bending_coeff = -bending_modulus per
 * vesicle and, when some vesicle has an excess density, the gravity traction
 * excess_density (g . x) n in both the right-hand side (of the positions)
 * and the matvec (of dt u) (src/InterfacialForce.cc:126-184).  A viscosity
 * contrast lambda != 1 on some vesicle adds the double layer of
 * (1 - lambda) u to the matvec's Stokes evaluation and scales its u by
 * (1 + lambda) / 2 (src/InterfacialVelocity.cc:945-968); the double-layer
 * self operators are then built at every step, and that step's far field
 * runs over the upsampled sources whatever far_coarse says.  Until called,
 * every vesicle has the step parameters' bending modulus, viscosity
 * contrast 1 and no excess density.  A negative viscosity contrast is
 * VES3D_INVALID_PARAMETER (nothing changes). */
int ves3d_hip_stepper_set_vesicle_props(void *handle, const double *bending_modulus_host,
                                        const double *viscosity_contrast_host,
                                        const double *excess_density_host,
                                        const double *gravity_field_host);
/* Overrides the recorded area / volume targets (host arrays [nves]). */
int ves3d_hip_stepper_set_area_volume(void *handle, const double *area_host,
                                      const double *vol_host);

/* ---- input formats (host only, no device work) --------------------------
 * DataIO::ReadDataStl in ASCII (src/DataIO_templates.cc:72-98): every number
 * of the file, '#' starting a comment to the end of its line.  With
 * out_host == NULL only counts. */
int ves3d_read_ascii(const char *path, double *out_host, size_t capacity, size_t *count_host);
/* ---- checkpoints (host only) ---------------------------------------------
 * The reference's ASCII checkpoint: Monitor writes Parameters::pack then
 * EvolveSurface::pack (src/Monitor.cc:58-78), Simulation reads them back to
 * restart (src/ves3d_simulation.cc:119-123, :237-249).  The parameters block
 * (Parameters::pack, src/Parameters.cc:424-470) as a struct; string-valued
 * parameters (file names) are written empty and skipped on reading. */
typedef struct {
  int version;                 /* "version:" of the blocks; >= 593 is the current layout */
  int n_surfs, sh_order, filter_freq, upsample_freq;
  double bending_modulus, viscosity_contrast, time_horizon, ts, time_tol;
  int time_iter_max, time_adaptive, solve_for_velocity, pseudospectral;
  char scheme[32], time_precond[32], bg_flow[32], singular_stokes[32], rep_type[32];
  int rep_maxit, rep_upsample, rep_filter_freq;
  double rep_ts, rep_tol, rep_exponent, repul_dist, bg_flow_param, periodic_length;
  int interaction_upsample, checkpoint;
  double checkpoint_stride, error_factor;
  int num_threads;
  double excess_density, gravity_field[3];
  /* the SURFACE block (Surface::pack, src/Surface.cc:609-627) */
  int sht_resample_order, diff_filter_freq, reparam_filter_freq;
} ves3d_checkpoint_params;
/* Parameters<T>::init defaults (src/Parameters.cc:23-58) for sh_order and
 * nves, with the globally implicit velocity scheme and the diagonal spectral
 * preconditioner. */
int ves3d_checkpoint_defaults(int sh_order, size_t nves, ves3d_checkpoint_params *prm);
/* Writes a checkpoint: positions X [nves][3][(p+1)2p] (host, AxisMajor) and
 * the vesicle properties [nves] (VesicleProperties::pack,
 * src/VesicleProps.cc:110-122).  Numbers in %.16e (Monitor's stream
 * settings), which round-trips every double. */
int ves3d_checkpoint_write(const char *path, const ves3d_checkpoint_params *prm, size_t nves,
                           const double *X, const double *bending_modulus,
                           const double *viscosity_contrast, const double *excess_density);
/* Reads a checkpoint written by the reference or by ves3d_checkpoint_write,
 * accepting the layouts of the reference's unpack methods (versions <= 590,
 * < 592 and < 593 differ).  X may be NULL (count only); property arrays may
 * be NULL; versions <= 590 store no properties and the parameters' values
 * are returned (VesicleProperties::setFromParams).  VES3D_IO_ERROR: the file
 * cannot be opened; VES3D_IO_BAD_STREAM: a key or block is not where the
 * format puts it. */
int ves3d_checkpoint_read(const char *path, ves3d_checkpoint_params *prm, size_t *nves_host,
                          double *X, size_t x_capacity, double *bending_modulus,
                          double *viscosity_contrast, double *excess_density,
                          size_t props_capacity);

/* The vesicle-properties file of Simulation::setup_from_options
 * (src/ves3d_simulation.cc:174-196): the first 3 nves numbers of the ASCII
 * file, one row (bending modulus, viscosity contrast, excess density) per
 * vesicle, into props_host [3][nves] ordered by property (the reference's
 * Transpose, VesicleProperties::getPropIdx order, src/VesicleProps.cc:61-72).
 * Fewer than 3 nves numbers: VES3D_IO_BAD_STREAM. */
int ves3d_read_vesicle_props(const char *path, size_t nves, double *props_host);

/* InitializeShapes (src/HelperFuns.cc:366-432): X_host [nves][3][(p+1)2p] from
 * a shape gallery (shapes [3][(p+1)2p] back to back, gallery_len values) and
 * a geometry specification of 8 values per vesicle: shape index, scale,
 * centre (3), rotation angles z, y, z (X = Rz Ry Rz (scale x) + centre). */
int ves3d_initialize_shapes(int p, size_t nves, const double *gallery, size_t gallery_len,
                            const double *geo_spec, double *X_host);

/* ---- the old pole-moving route (MovePole, include/MovePole.h:8-30) --------
 * Grid fields [nsub][(p+1)2p], FP64.  scheme: 0 Direct, 1 ViaSpHarm,
 * 2 DirectEagerEval (enum SingularStokesRot).  all_rot_mats_host /
 * sh_rot_mats_host: the reference's precomputed/all_rot_mats_{p}_double.txt
 * and SpHarmRotMats_p{p}_double.txt contents (host memory, file order), or
 * NULL to build them from their definition (any p).  Replaces the MovePole
 * constructor (src/MovePole.cc:1-95). */
int ves3d_hip_movepole_create(int p, int scheme, const double *all_rot_mats_host,
                              const double *sh_rot_mats_host, void **handle_host);
int ves3d_hip_movepole_destroy(void *handle);
/* MovePole::setOperands (src/MovePole.cc:105-156): num operands, operand v
 * a device array of nsub_host[v] grid functions; the handle keeps the
 * pointers (valid until the next call). */
int ves3d_hip_movepole_set_operands(void *handle, const double *const *arr_host,
                                    const size_t *nsub_host, int num, ves3d_stream_t stream);
/* MovePole::operator() (src/MovePole.cc:158-163): results_host[v] (device
 * arrays shaped like operand v) = operand v with its pole moved to grid
 * point (trg_i, trg_j). */
int ves3d_hip_movepole_apply(void *handle, int trg_i, int trg_j, double *const *results_host);
/* Copies the handle's direct rotation matrices to the host,
 * [p+1][(p+1)2p][(p+1)2p] row-major (row = output point). */
int ves3d_hip_movepole_matrices(void *handle, double *all_rot_mats_host);
/* Device<>::AggregateRotation (src/DeviceCPU.cc:644-681): for jj < nlong and
 * ii < n_vec, res_host[n_vec jj + ii] = CircShift(mat, jj np) vec_host[ii],
 * mat the column-major np x np matrix of one latitude (np = nlat nlong).
 * The reference's wrk / n_stream arguments are internal here.  Synchronous. */
int ves3d_hip_aggregate_rotation_f64(int sh_order, int nlat, int nlong, int n_vec,
                                     const int *n_sub_host, const double *mat,
                                     const double *const *vec_host, double *const *res_host,
                                     ves3d_stream_t stream);

/* ---- near-singular correction (NearSingular<Real>, include/NearSingular.h:8-30)
 * For surface targets (trg_is_surf) and a single-layer density, FP64.
 * Sources: the upsampled surfaces with the two poles first, point-major
 * [nves][M][3], M = 2 + 2 p_up (p_up + 1) (scoord_far,
 * src/StokesVelocity.cc:199-218).  Targets: nves blocks of (p+1)2p points,
 * point-major [nves (p+1)2p][3] (tcoord_repl, :227-239).  A near pair is a
 * target and a vesicle other than its own with a source point closer than
 * r_near, paired with the closest such point (src/NearSingular.cc:797-918).
 * The handle keeps pointers to the coordinates passed to near_set_coord:
 * they must stay valid until the next near_set_coord. */
int ves3d_hip_near_create(int p, int p_up, size_t nves, double repul_dist, double r_near,
                          void **handle_host);
int ves3d_hip_near_destroy(void *handle);
/* SetSrcCoord + SetTrgCoord + SetupCoordData (src/NearSingular.cc:40-80,
 * :80-1120): near-pair search, patch projection, repulsion force.
 * Synchronises the stream. */
int ves3d_hip_near_set_coord(void *handle, const double *src, const double *trg,
                             ves3d_stream_t stream);
/* Number of near pairs and the near-zone radius of the last set_coord. */
int ves3d_hip_near_info(void *handle, size_t *npairs_host, double *r_near_host);
/* ForceRepul (:1122-1126), added to f [nves][3][(p+1)2p] (AxisMajor). */
int ves3d_hip_near_force_repul(void *handle, double *f, ves3d_stream_t stream);
/* operator() minus SubtractDirect (:1143-1212, :1375-1551): vel
 * [nves][3][(p+1)2p] += near-singular velocity - direct sum, for every
 * near target.  q: SL density times area times quadrature weight,
 * point-major [nves][M][3] (qforce_single, src/StokesVelocity.cc:308-334);
 * vup [nves][3][2 p_up (p_up+1)] and vpole [nves][3][2]: the surface
 * velocity on the upsampled grid and at the poles (S_vel_up, :452-470). */
int ves3d_hip_near_apply(void *handle, const double *q, const double *vup, const double *vpole,
                         double *vel, ves3d_stream_t stream);

/* ---- learned SH-space preconditioner ------------------------------------
 * No counterpart in the reference: a per-vesicle network that fills the
 * preconditioner slot of ImplicitPrecond (src/InterfacialVelocity.cc:1054-1109).
 * For the Galerkin vector r, with D the diagonal spectral preconditioner and
 * z_v = (D r)_v the 4 p(p+2) coefficients of vesicle v, s_v = |z_v|, and
 * F_v its canonical frame (ves3d_hip_stepper_frames) acting on the three
 * velocity blocks:
 *     P(r)_v = z_v + s_v K F_v g_o(F_v^T z_v / s_v),
 * g_o(z) = (g(z) - g(-z)) / 2 the odd part of the network, so that
 * P(a r) = a P(r) for every real a,
 * K zeroes the tension coefficients of degree > filter_freq: Surface::div
 * is low-pass filtered at that degree (src/Surface.cc:372), so they span the
 * null space of the implicit operator, and keeping the correction out of them
 * makes the solve reach the velocity of the diagonal-preconditioned one.
 * g: nin -> width -> width -> nin with tanh-GELU, FP32 weights, parameters
 * laid out [W1 (width x nin), b1, W2 (width x width), b2, W3 (nin x width), b3].
 * Rotating a shape rotates its own operator's velocity blocks by the same
 * matrix, so one network serves vesicles in every orientation.  Trained
 * online on (A_loc e, e) pairs recorded during solves: e the solver's own
 * Krylov vectors, A_loc the vesicles' own operators
 * (ves3d_hip_stepper_matvec_local), mapped into the canonical frames. */
int ves3d_hip_lpc_create(int nin, int width, size_t capacity, unsigned long long seed,
                         void **handle_host);
int ves3d_hip_lpc_destroy(void *handle);
int ves3d_hip_lpc_num_params(void *handle, size_t *n_host);
/* Number of training samples recorded so far (the buffer keeps the last
 * `capacity`). */
int ves3d_hip_lpc_samples(void *handle, size_t *n_host);
/* Device-to-device copies of the FP32 parameter vector; setting parameters
 * also zeroes the Adam moments and step count. */
int ves3d_hip_lpc_set_params(void *handle, const float *params);
int ves3d_hip_lpc_get_params(void *handle, float *params);
/* bf16 != 0: every product of the forward and backward passes takes BF16
 * operands (the activations, and the weights rounded to nearest even from
 * the FP32 master copy after each update) on BF16 MFMA tiles
 * (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; pre-activations,
 * biases, gradients and Adam stay FP32.  0 (the default): FP32 activations
 * on FP32 MFMA tiles.  BF16 needs nin and width multiples of 8
 * (VES3D_INVALID_PARAMETER otherwise). */
int ves3d_hip_lpc_set_activations(void *handle, int bf16);
/* Y [B][nin] = g(Z [B][nin]). */
int ves3d_hip_lpc_forward(void *handle, const float *Z, size_t B, float *Y,
                          ves3d_stream_t stream);
/* loss = sum((g(Z) - T)^2) / (2 B nin) into *loss_host (may be NULL); the
 * gradient with respect to the parameters into grads (device, may be NULL). */
int ves3d_hip_lpc_loss_grad(void *handle, const float *Z, const float *T, size_t B,
                            double *loss_host, float *grads, ves3d_stream_t stream);
/* Decoupled weight decay of the Adam steps (AdamW; default 0, plain Adam). */
int ves3d_hip_lpc_set_weight_decay(void *handle, float wd);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
int ves3d_hip_lpc_adam(void *handle, float lr, ves3d_stream_t stream);
/* epochs passes over nine tenths of the recorded samples in shuffled
 * minibatches of `batch` (each sample enters with both signs), Adam with a
 * cosine decay of lr; keeps the parameters with the lowest loss on the
 * held-out tenth (evaluated after every epoch) and returns that loss. */
int ves3d_hip_lpc_train(void *handle, int epochs, int batch, float lr, double *last_loss_host);
/* Loss of the odd part of the network over all recorded samples and the mean
 * square of their targets (their ratio is the network's relative fit). */
int ves3d_hip_lpc_dataset_loss(void *handle, double *loss_host, double *target_ms_host);
/* Attaches a learned preconditioner to a stepper (precond = 2 uses it once
 * it has been trained or given parameters); collect != 0 records training
 * pairs during every solve. */
int ves3d_hip_stepper_attach_pc(void *stepper, void *pc, int collect);

#ifdef __cplusplus
}
#endif
#endif /* VES3D_HIP_H */

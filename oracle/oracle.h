/*
 * oracle.h -- FP64 CPU restatement of the Ves3D per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in ves3d_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product path never links, loads or calls anything here.
 *
 * Every function restates a reference routine (lu1and10/ves3d) and cites the
 * file:line it follows.  Layouts are the reference's:
 *   - per-vesicle fields are AxisMajor [ves][xyz][stride]      (include/Vectors.h)
 *   - all-pairs / far-field arrays are point-major [pt][xyz]    (src/CPUKernels.cc:966)
 *   - SH coefficients in the SphericalHarmonics layout are function-major,
 *     p(p+2) values per function (src/SphericalHarmonics.cc:30-40).
 *
 * Parity pins (what checks each group; see tests/test_oracle_*.py):
 *   direct sums      : the reference's own src/CPUKernels.cc compiled into
 *                      oracle/_ref (bitwise-close), closed forms on the sphere.
 *   Gauss-Legendre   : numpy.polynomial.legendre.leggauss, exactness on
 *                      polynomials, precomputed/quad_weights_{6,12} files.
 *   SHT              : precomputed/legTrans*{6,12}_double.txt fixtures, round
 *                      trips, closed-form Y_1m, scipy lpmv.
 *   rotation         : orthogonality per degree, rotated-evaluation identity.
 *   singular quad.   : closed forms on the unit sphere (SL[c]=2/3 c, SL[n]=0,
 *                      DL[c]=-1/2 c), convergence in the upsampling order.
 */
#ifndef VES3D_ORACLE_H
#define VES3D_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- direct Stokes sums (direct.c) ---------------------------------- */
void orc_direct_stokes(int src_stride, int trg_stride, int n_surfs,
                       int trg_idx_head, int trg_idx_tail, const double *qw,
                       const double *trg, const double *src, const double *den,
                       double *pot);
void orc_direct_stokes_dl(int src_stride, int trg_stride, int n_surfs,
                          int trg_idx_head, int trg_idx_tail, const double *qw,
                          const double *trg, const double *src,
                          const double *norm, const double *den, double *pot);
void orc_stokes_all2all(const double *src, const double *den, size_t np,
                        double *pot);
void orc_stokes_sl_dl_eval(const double *src_pos, const double *sl_den,
                           const double *dl_den, size_t n_src,
                           const double *trg_pos, double *trg_vel, size_t n_trg);
void orc_far_field(const double *src_pos, const double *sl_den,
                   const double *dl_den, size_t n_src, const double *trg_pos,
                   double *trg_vel, size_t n_trg, size_t spg, size_t tpg);
void orc_stokes_all2all_f32(const float *src, const float *den, size_t np,
                            float *pot);

#ifdef __cplusplus
}
#endif
#endif

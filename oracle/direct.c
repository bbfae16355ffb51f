/*
 * direct.c -- oracle restatement of the reference's direct Stokes sums.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *  orc_direct_stokes        <- src/CPUKernels.cc:8-118 (DirectStokesKernel,
 *                              DirectStokesKernel_Noqw); Device<CPU>::DirectStokes
 *                              src/DeviceCPU.cc:513-541.
 *  orc_direct_stokes_dl     <- src/CPUKernels.cc:1059-1136
 *                              (DirectStokesDoubleLayerKernel_Template).
 *  orc_stokes_all2all       <- src/CPUKernels.cc:966-1054 (StokesAlltoAll).
 *  orc_stokes_sl_dl_eval    <- the direct-sum stand-in for PVFMMEval
 *                              (include/PVFMMInterface.h:53) with the pair kernels
 *                              stokes_sl_uKernel / stokes_dl_uKernel
 *                              (src/PVFMMInterface.cc:161-238, :284-369):
 *                              d = trg - src, pairs with r^2 <= 128*eps skipped,
 *                              SL (1/8pi)[f/r + (f.d)d/r^3],
 *                              DL (3/4pi)(d.n)(d.f)d/r^5.
 *                              dl_den per source is [f*w (3), n (3)], the packing
 *                              of qforce_double in src/StokesVelocity.cc:351-372.
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include "oracle.h"

#define ORC_PI 3.14159265358979323846

void orc_direct_stokes(int src_stride, int trg_stride, int n_surfs,
                       int trg_idx_head, int trg_idx_tail, const double *qw,
                       const double *trg, const double *src, const double *den,
                       double *pot)
{
    const double inv8pi = 1.0 / ORC_PI / 8.0;
#pragma omp parallel for schedule(static)
    for (int vt = 0; vt < n_surfs; vt++) {
        const double *S = src + (size_t)3 * src_stride * vt;
        const double *D = den + (size_t)3 * src_stride * vt;
        const double *T = trg + (size_t)3 * trg_stride * vt;
        double *P = pot + (size_t)3 * trg_stride * vt;
        for (int t = trg_idx_head; t < trg_idx_tail; t++) {
            double tx = T[t], ty = T[trg_stride + t], tz = T[2 * trg_stride + t];
            double px = 0, py = 0, pz = 0;
            for (int s = 0; s < src_stride; s++) {
                double dx = S[s] - tx;
                double dy = S[src_stride + s] - ty;
                double dz = S[2 * src_stride + s] - tz;
                double ir = dx * dx + dy * dy + dz * dz;
                if (ir != 0) ir = 1.0 / sqrt(ir);
                double w = qw ? qw[s] : 1.0;
                double fx = D[s] * w, fy = D[src_stride + s] * w,
                       fz = D[2 * src_stride + s] * w;
                double cc = (dx * fx + dy * fy + dz * fz) * ir * ir;
                px += (fx + cc * dx) * ir;
                py += (fy + cc * dy) * ir;
                pz += (fz + cc * dz) * ir;
            }
            P[t] = px * inv8pi;
            P[trg_stride + t] = py * inv8pi;
            P[2 * trg_stride + t] = pz * inv8pi;
        }
    }
}

void orc_direct_stokes_dl(int src_stride, int trg_stride, int n_surfs,
                          int trg_idx_head, int trg_idx_tail, const double *qw,
                          const double *trg, const double *src,
                          const double *norm, const double *den, double *pot)
{
    const double scal = -3.0 / (4.0 * ORC_PI);
#pragma omp parallel for schedule(static)
    for (int vt = 0; vt < n_surfs; vt++) {
        const double *S = src + (size_t)3 * src_stride * vt;
        const double *N = norm + (size_t)3 * src_stride * vt;
        const double *D = den + (size_t)3 * src_stride * vt;
        const double *T = trg + (size_t)3 * trg_stride * vt;
        double *P = pot + (size_t)3 * trg_stride * vt;
        for (int t = trg_idx_head; t < trg_idx_tail; t++) {
            double tx = T[t], ty = T[trg_stride + t], tz = T[2 * trg_stride + t];
            double px = 0, py = 0, pz = 0;
            for (int s = 0; s < src_stride; s++) {
                double dx = S[s] - tx;
                double dy = S[src_stride + s] - ty;
                double dz = S[2 * src_stride + s] - tz;
                double ir = dx * dx + dy * dy + dz * dz;
                if (ir != 0) ir = 1.0 / sqrt(ir);
                double ir5 = ir * ir;
                ir5 = ir5 * ir5 * ir;
                double rn = N[s] * dx + N[src_stride + s] * dy + N[2 * src_stride + s] * dz;
                double rf = D[s] * dx + D[src_stride + s] * dy + D[2 * src_stride + s] * dz;
                double c = rn * rf * ir5;
                if (qw) c *= qw[s];
                px += dx * c;
                py += dy * c;
                pz += dz * c;
            }
            P[t] = px * scal;
            P[trg_stride + t] = py * scal;
            P[2 * trg_stride + t] = pz * scal;
        }
    }
}

void orc_stokes_all2all(const double *src, const double *den, size_t np,
                        double *pot)
{
    const double inv8pi = 1.0 / ORC_PI / 8.0;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)np; t++) {
        double tx = src[3 * t], ty = src[3 * t + 1], tz = src[3 * t + 2];
        double px = 0, py = 0, pz = 0;
        for (size_t s = 0; s < np; s++) {
            double dx = src[3 * s] - tx, dy = src[3 * s + 1] - ty,
                   dz = src[3 * s + 2] - tz;
            double ir = dx * dx + dy * dy + dz * dz;
            if (ir != 0) ir = 1.0 / sqrt(ir);
            double fx = den[3 * s], fy = den[3 * s + 1], fz = den[3 * s + 2];
            double cc = (dx * fx + dy * fy + dz * fz) * ir * ir;
            px += (fx + cc * dx) * ir;
            py += (fy + cc * dy) * ir;
            pz += (fz + cc * dz) * ir;
        }
        pot[3 * t] = px * inv8pi;
        pot[3 * t + 1] = py * inv8pi;
        pot[3 * t + 2] = pz * inv8pi;
    }
}

/* Single-precision all-pairs sum, same arithmetic order as the float
 * StokesAlltoAll (src/CPUKernels.cc:966-1008); used as the FP32 CPU baseline. */
void orc_stokes_all2all_f32(const float *src, const float *den, size_t np,
                            float *pot)
{
    const float inv8pi = (float)(1.0 / ORC_PI / 8.0);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)np; t++) {
        float tx = src[3 * t], ty = src[3 * t + 1], tz = src[3 * t + 2];
        float px = 0, py = 0, pz = 0;
        for (size_t s = 0; s < np; s++) {
            float dx = src[3 * s] - tx, dy = src[3 * s + 1] - ty,
                  dz = src[3 * s + 2] - tz;
            float ir = dx * dx + dy * dy + dz * dz;
            if (ir != 0) ir = 1.0f / sqrtf(ir);
            float fx = den[3 * s], fy = den[3 * s + 1], fz = den[3 * s + 2];
            float cc = (dx * fx + dy * fy + dz * fz) * ir * ir;
            px += (fx + cc * dx) * ir;
            py += (fy + cc * dy) * ir;
            pz += (fz + cc * dz) * ir;
        }
        pot[3 * t] = px * inv8pi;
        pot[3 * t + 1] = py * inv8pi;
        pot[3 * t + 2] = pz * inv8pi;
    }
}

void orc_stokes_sl_dl_eval(const double *src_pos, const double *sl_den,
                           const double *dl_den, size_t n_src,
                           const double *trg_pos, double *trg_vel, size_t n_trg)
{
    orc_far_field(src_pos, sl_den, dl_den, n_src, trg_pos, trg_vel, n_trg, 0, 0);
}

/* The same sum with same-group (same-vesicle) pairs left out when spg, tpg > 0:
 * what remains of PVFMMEval after NearSingular::SubtractDirect removes the
 * target's own vesicle (src/NearSingular.cc:1143-1212, trg_is_surf). */
void orc_far_field(const double *src_pos, const double *sl_den,
                   const double *dl_den, size_t n_src, const double *trg_pos,
                   double *trg_vel, size_t n_trg, size_t spg, size_t tpg)
{
    const double eps = DBL_EPSILON * 128;
    const double csl = 1.0 / (8.0 * ORC_PI);
    const double cdl = 3.0 / (4.0 * ORC_PI);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)n_trg; t++) {
        double tx = trg_pos[3 * t], ty = trg_pos[3 * t + 1], tz = trg_pos[3 * t + 2];
        double sx = 0, sy = 0, sz = 0, qx = 0, qy = 0, qz = 0;
        long tg = (spg && tpg) ? t / (long)tpg : -1;
        for (size_t s = 0; s < n_src; s++) {
            if (tg >= 0 && (long)(s / spg) == tg) continue;
            double dx = tx - src_pos[3 * s], dy = ty - src_pos[3 * s + 1],
                   dz = tz - src_pos[3 * s + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > eps)) continue;
            double ir = 1.0 / sqrt(r2);
            double ir2 = ir * ir;
            if (sl_den) {
                const double *f = sl_den + 3 * s;
                double ip = (f[0] * dx + f[1] * dy + f[2] * dz) * ir2;
                sx += ir * (f[0] + dx * ip);
                sy += ir * (f[1] + dy * ip);
                sz += ir * (f[2] + dz * ip);
            }
            if (dl_den) {
                const double *f = dl_den + 6 * s, *n = dl_den + 6 * s + 3;
                double ir5 = ir2 * ir2 * ir;
                double p = (n[0] * dx + n[1] * dy + n[2] * dz) *
                           (f[0] * dx + f[1] * dy + f[2] * dz) * ir5;
                qx += dx * p;
                qy += dy * p;
                qz += dz * p;
            }
        }
        trg_vel[3 * t] = sx * csl + qx * cdl;
        trg_vel[3 * t + 1] = sy * csl + qy * cdl;
        trg_vel[3 * t + 2] = sz * csl + qz * cdl;
    }
}

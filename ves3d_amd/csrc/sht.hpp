// sht.hpp -- the SphericalHarmonics transforms on the GPU
// (SHC2Grid, Grid2SHC, SHC2GridTranspose, SHC2Pole; src/SphericalHarmonics.cc:3-360)
// expressed as two grouped GEMM launches each (Legendre stage per order,
// Fourier stage as one dense product), with all re-layouts folded into the
// operand strides.
#pragma once
#include "gemm.hpp"
#include "sh_tables.hpp"

namespace v3d {

// Workspace element counts the transforms need.
inline long shc2grid_work(long N, int p0, int p1) { return 2L * p0 * N * (p1 + 1); }
inline long grid2shc_work(long N, int p0, int p1) { return 2L * p1 * N * (p0 + 1); }
// scratch of sht_synth_multi with nvar variants
inline long shc2grid_multi_work(long N, int p0, int p1, int nvar) {
  return nvar * shc2grid_work(N, p0, p1);
}

// S [N][p0(p0+2)] -> X [N][(p1+1)2p1]; X_theta / X_phi optional (may be null).
template <typename T>
int sht_shc2grid(const T *S, long N, int p0, int p1, T *X, T *X_theta, T *X_phi,
                 T *work, hipStream_t st);
// X [N][(p0+1)2p0] -> S [N][p1(p1+2)].
template <typename T>
int sht_grid2shc(const T *X, long N, int p0, int p1, T *S, T *work, hipStream_t st);
// Adjoint of SHC2Grid(p1 -> p0): X [N][(p0+1)2p0] -> S [N][p1(p1+2)].
template <typename T>
int sht_shc2grid_transpose(const T *X, long N, int p0, int p1, T *S, T *work,
                           hipStream_t st);
// Synthesis with the latitude (Legendre) stage differentiated dleg times and
// the longitude (Fourier) stage dfour times, dleg, dfour in {0,1,2}; the
// Legendre stage is scaled by alpha.  Same workspace as sht_shc2grid.
template <typename T>
int sht_synth(const T *S, long N, int p0, int p1, int dleg, int dfour, T *X, T *work,
              hipStream_t st, T alpha);
// S [N][p0(p0+2)] -> P [N][2] at x = -1, +1.
template <typename T>
int sht_shc2pole(const T *S, long N, int p0, T *P, hipStream_t st);

// Several syntheses of the same coefficients S [N][p0(p0+2)] in two grouped
// launches (one Legendre, one Fourier stage) instead of two per variant:
// variant v takes dleg[v] latitude and dfour[v] longitude derivatives
// (0..2; neg[v] with dleg[v] = 1 negates the latitude derivative, SHTrans's
// sign) and writes X + v x_stride, [N][(p1+1) 2p1].  nvar <= 4; work:
// shc2grid_multi_work(N, p0, p1, nvar) elements.
template <typename T>
int sht_synth_multi(const T *S, long N, int p0, int p1, int nvar, const int *dleg,
                    const int *dfour, const int *neg, T *X, long x_stride, T *work,
                    hipStream_t st);

// The matrices of SHC2GridTranspose (p1 -> p0) on the device, for a caller
// that runs its stages itself: sht_mft gives MatFourier(p1, p0) ([2p1][2p0]:
// frequency row jj, longitude column), sht_mlt MatLegendre(p1, p0)[m] for
// m = 0..p1, consecutive blocks [(p1+1-m)][p0+1].
template <typename T>
const T *sht_mft(int p0, int p1);
template <typename T>
const T *sht_mlt(int p0, int p1);

}  // namespace v3d

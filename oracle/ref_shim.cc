// ref_shim.cc -- extern "C" entry points over the reference's own CPU Stokes
// kernels, so tests can call them through ctypes.  TEST INFRASTRUCTURE ONLY.
// Compiled together with /root/reference/src/CPUKernels.cc by build_ref.sh;
// this file only forwards arguments.  Declarations: include/CPUKernels.h:9-60.
#include "CPUKernels.h"

extern "C" {
void ref_direct_stokes_f64(int ss, int ts, int n, int h, int t, const double *qw,
                           const double *trg, const double *src,
                           const double *den, double *pot) {
  if (qw) DirectStokesKernel(ss, ts, n, h, t, qw, trg, src, den, pot);
  else DirectStokesKernel_Noqw(ss, ts, n, h, t, trg, src, den, pot);
}
void ref_direct_stokes_sse_f64(int ss, int ts, int n, int h, int t,
                               const double *qw, const double *trg,
                               const double *src, const double *den,
                               double *pot) {
  if (qw) DirectStokesSSE(ss, ts, n, h, t, qw, trg, src, den, pot);
  else DirectStokesSSE_Noqw(ss, ts, n, h, t, trg, src, den, pot);
}
void ref_direct_stokes_f32(int ss, int ts, int n, int h, int t, const float *qw,
                           const float *trg, const float *src, const float *den,
                           float *pot) {
  if (qw) DirectStokesKernel(ss, ts, n, h, t, qw, trg, src, den, pot);
  else DirectStokesKernel_Noqw(ss, ts, n, h, t, trg, src, den, pot);
}
void ref_direct_stokes_dl_f64(int ss, int ts, int n, int h, int t,
                              const double *qw, const double *trg,
                              const double *src, const double *norm,
                              const double *den, double *pot) {
  DirectStokesDoubleLayerKernel(ss, ts, n, h, t, qw, trg, src, norm, den, pot);
}
void ref_stokes_all2all_f64(const double *src, const double *den, size_t np,
                            double *pot) {
  StokesAlltoAll(src, den, np, pot, 0);
}
void ref_stokes_all2all_f32(const float *src, const float *den, size_t np,
                            float *pot) {
  StokesAlltoAll(src, den, np, pot, 0);
}
}

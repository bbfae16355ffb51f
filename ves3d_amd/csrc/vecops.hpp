// vecops.hpp -- BLAS-1 style device helpers (Device<>::axpy / AlgebraicDot /
// MaxAbs counterparts) used by the solver and the implicit step.
#pragma once
#include "common.hpp"

namespace v3d {

// y = a*x + b*y  (x may alias y; b == 0 writes a*x without reading y)
int dev_axpby(long n, double a, const double *x, double b, double *y, hipStream_t st);
// z = a*x + y
int dev_axpy3(long n, double a, const double *x, const double *y, double *z, hipStream_t st);
int dev_copy(long n, const double *x, double *y, hipStream_t st);
int dev_fill(long n, double v, double *y, hipStream_t st);
// Several dot products in one pass: out[k] = sum_i x_k[i] y_k[i], k < nd <= 4,
// accumulated in FP64 with a fixed reduction order (deterministic).  The
// results land in host memory `out_host` (synchronises the stream).
int dev_dots(long n, int nd, const double *const *x, const double *const *y,
             double *out_host, hipStream_t st);
double dev_norm(long n, const double *x, hipStream_t st, int *err);

// Vectors V[k] = V + k ldv, k < nk <= kMaxVecCoefs (an Arnoldi basis).
constexpr int kMaxVecCoefs = 64;
struct VecCoefs {
  double c[kMaxVecCoefs];
};
// out[k] = V[k] . w for k < nk, fixed-order FP64 reductions, into host
// memory (synchronises the stream).
int dev_mdots(long n, int nk, const double *V, long ldv, const double *w, double *out_host,
              hipStream_t st);
// w = b w + sum_k c[k] V[k] (b == 0: w is not read)
int dev_combine(long n, int nk, const double *V, long ldv, const double *c_host, double b,
                double *w, hipStream_t st);

}  // namespace v3d

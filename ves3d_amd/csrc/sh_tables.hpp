// sh_tables.hpp -- host-side construction (FP64) of the spherical-harmonic
// matrices of the reference's SphericalHarmonics<Real> class
// (src/SphericalHarmonics.cc:543-860) and the device-resident copies the
// kernels read.  Built once per (p0, p1) and cached for the process lifetime.
#pragma once
#include <vector>

namespace v3d {

// Gauss-Legendre nodes (ascending) and weights on [-1, 1] with n points
// (the rule cgqf computes, src/SphericalHarmonics.cc:549-551).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);

// LegPoly (src/SphericalHarmonics.cc:888-928): out[(deg+1)(deg+2)/2][N].
void legpoly(const double *X, long N, int degree, double *out);
// LegPolyDeriv (src/SphericalHarmonics.cc:930-950).
void legpoly_deriv(const double *X, long N, int degree, double *out);
// Second latitude derivative by the raising/lowering ladder the reference's
// d2legTrans matrices come from (matlab/Ynm.m:27-40), LegPoly row order.
void legpoly_d2(const double *X, long N, int degree, double *out);

inline int sh_ncoef(int p) { return p * (p + 2); }
inline int sh_ngrid(int p) { return 2 * p * (p + 1); }
inline int sh_blocklen(int p, int j) { return p + 1 - (j + 1) / 2; }
// Offset of coefficient block j (0..2p-1) inside one function.
inline int sh_blockoff(int p, int j) {
  int o = 0;
  for (int i = 0; i < j; ++i) o += sh_blocklen(p, i);
  return o;
}

// All matrices for one (p0, p1) pair, row-major, concatenated in one array.
struct ShMatsHost {
  int p0, p1;
  std::vector<double> data;
  long Mf, Mdf, Mfinv;          // MatFourier(p0,p1), MatFourierGrad(p0,p1), MatFourierInv(p0,p1): [2p0][2p1]
  long MfT;                     // MatFourier(p1,p0): [2p1][2p0] (read transposed by SHC2GridTranspose)
  long Md2f;                    // second longitude derivative of MatFourier: [2p0][2p1]
  std::vector<long> Ml, Mdl;    // MatLegendre(p0,p1)[m], MatLegendreGrad: [(p0+1-m)][(p1+1)], m=0..p0
  std::vector<long> Mdln;       // -MatLegendreGrad (SHTrans's first latitude derivative), same shape
  std::vector<long> Md2l;       // second latitude derivative (ladder of d2legTrans), same shape
  std::vector<long> Mlinv;      // MatLegendreInv(p0,p1)[m]: [(p0+1)][(p1+1-m)], m=0..p1
  std::vector<long> MlT;        // MatLegendre(p1,p0)[m]: [(p1+1-m)][(p0+1)], m=0..p1
};
void build_sh_mats(int p0, int p1, ShMatsHost &h);

// Per-order rotation blocks (MatRotate, src/SphericalHarmonics.cc:759-860):
// block (l, n) at offset off[l*(p+1)+n], size len(n)^2, len(n) = min(2n+1, 2p).
struct RotMatsHost {
  int p;
  std::vector<double> data;
  std::vector<long> off;
  std::vector<int> len;
  std::vector<int> perm;      // degree-major position -> coefficient index (coeff_perm)
  std::vector<int> doff;      // degree-major offset of degree n
};
void build_rot_mats(int p, RotMatsHost &h);

// LegendreWeights(p) and SingularWeights(p) (src/SphericalHarmonics.cc:557-604).
void legendre_weights(int p, std::vector<double> &w);
void singular_weights(int p, std::vector<double> &w);

// Host FP64 Grid2SHC used while building MatRotate.
void host_grid2shc(const ShMatsHost &m, const double *X, long N, double *S);

// Device-resident copies.
struct ShMatsDev {
  int p0 = 0, p1 = 0;
  const ShMatsHost *h = nullptr;
  double *d64 = nullptr;
  float *d32 = nullptr;
};
const ShMatsDev *sh_mats(int p0, int p1);  // nullptr on allocation failure

struct RotMatsDev {
  const RotMatsHost *h = nullptr;
  double *d64 = nullptr;
  float *d32 = nullptr;
  int *perm = nullptr;     // degree-major position -> coefficient index
  int *iperm = nullptr;    // coefficient index -> degree-major position
  int *partner = nullptr;  // coefficient index -> cos/sin partner coefficient index (or itself)
  int *order = nullptr;    // coefficient index -> order m
  int *is_sin = nullptr;   // coefficient index -> 1 if sine coefficient
  double *phase = nullptr; // [p][p+1][2]: cos, sin of k j pi / p (longitude j, order k)
  // per coefficient c of RotateTranspose's output: degree-major sources of
  // c and of its cos/sin partner, its order, and 0 (m = 0 or p: a copy),
  // 1 (cosine) or 2 (sine) -- [Ncoef][4]
  int *scatter = nullptr;
};
const RotMatsDev *rot_mats(int p);

}  // namespace v3d

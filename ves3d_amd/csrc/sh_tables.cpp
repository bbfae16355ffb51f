// sh_tables.cpp -- host construction of the SphericalHarmonics matrices.
// Formulas follow src/SphericalHarmonics.cc (line numbers at each function).
#include "sh_tables.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <memory>
#include <mutex>

namespace v3d {

static const double kPI = 3.14159265358979323846;

void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    // Newton on P_n from the Chebyshev-like initial guess; node i descending.
    double z = std::cos(kPI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p0 = 1.0;
        p1 = z;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) {
      p0 = 1.0;
      p1 = z;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;  // ascending
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// src/SphericalHarmonics.cc:888-928
void legpoly(const double *X, long N, int degree, double *out) {
  std::vector<double> u(N);
  const double fact = 1.0 / std::sqrt(4 * kPI);
  for (long n = 0; n < N; ++n) {
    u[n] = (X[n] * X[n] > 1.0) ? 0.0 : std::sqrt(1 - X[n] * X[n]);
    out[n] = fact;
  }
  std::vector<long> start(degree + 1, 0);
  for (int m = 1; m <= degree; ++m) start[m] = start[m - 1] + (degree - m + 2);
  for (int i = 1; i <= degree; ++i) {
    double c = (i == 1) ? std::sqrt(3.0 / 2.0) : 1.0;
    if (i > 1) c *= std::sqrt((double)(2 * i + 1) / (2 * i));
    for (long n = 0; n < N; ++n) out[N * start[i] + n] = -out[N * start[i - 1] + n] * u[n] * c;
  }
  for (int m = 0; m < degree; ++m) {
    double *pv = out + N * start[m];
    for (long n = 0; n < N; ++n) {
      double pmm = 0, pmmp1 = pv[n];
      for (int ll = m + 1; ll <= degree; ++ll) {
        double a = std::sqrt(((double)(2 * ll - 1) * (2 * ll + 1)) / ((ll - m) * (ll + m)));
        double b = std::sqrt(((double)(2 * ll + 1) * (ll + m - 1) * (ll - m - 1)) /
                             ((double)(ll - m) * (ll + m) * (2 * ll - 3)));
        double pll = X[n] * a * pmmp1 - b * pmm;
        pmm = pmmp1;
        pmmp1 = pll;
        pv[N * (ll - m) + n] = pll;
      }
    }
  }
}

// src/SphericalHarmonics.cc:930-950
void legpoly_deriv(const double *X, long N, int degree, double *out) {
  std::vector<double> lp((long)(degree + 1) * (degree + 2) / 2 * N);
  legpoly(X, N, degree, lp.data());
  auto idx = [degree](int m, int n) { return (long)((degree * 2 - m + 1) * m) / 2 + n; };
  for (int m = 0; m <= degree; ++m)
    for (int n = m; n <= degree; ++n) {
      const double *Pn = &lp[N * idx(m, n)];
      const double *Pn_ = (m + 1 <= n) ? &lp[N * idx(m + 1, n)] : &lp[0];
      double c1 = m;
      double c2 = (m + 1 <= n) ? std::sqrt((double)(n + m + 1)) * std::sqrt((double)(n > m ? n - m : 1)) : 0.0;
      double *H = out + N * idx(m, n);
      for (long i = 0; i < N; ++i) {
        double s = std::sqrt(1 - X[i] * X[i]);
        H[i] = -(c1 * X[i] * Pn[i] + c2 * s * Pn_[i]) / s;
      }
    }
}

// D P_m = c_m P_{m-1} - c_{m+1} P_{m+1}, c_k = sqrt((n+k)(n+1-k))/2, applied
// twice, with P_{-m} = (-1)^m P_m (orthonormal functions, Condon-Shortley).
void legpoly_d2(const double *X, long N, int degree, double *out) {
  std::vector<double> lp((long)(degree + 1) * (degree + 2) / 2 * N);
  legpoly(X, N, degree, lp.data());
  auto idx = [degree](int m, int n) { return (long)((degree * 2 - m + 1) * m) / 2 + n; };
  std::vector<double> full(2 * degree + 5), dp(2 * degree + 5);
  for (long i = 0; i < N; ++i)
    for (int n = 0; n <= degree; ++n) {
      // index k + (n+2) holds order k, k in [-n-2, n+2]
      auto at = [n](int k) { return k + n + 2; };
      auto c = [n](int k) {
        double v = (double)(n + k) * (n + 1 - k);
        return v > 0 ? std::sqrt(v) / 2.0 : 0.0;
      };
      for (int k = -n - 2; k <= n + 2; ++k) {
        double v = 0;
        if (k >= 0 && k <= n) v = lp[N * idx(k, n) + i];
        else if (k < 0 && -k <= n) v = ((-k) % 2 ? -1.0 : 1.0) * lp[N * idx(-k, n) + i];
        full[at(k)] = v;
      }
      for (int k = -n - 2; k <= n + 2; ++k) dp[at(k)] = 0;
      for (int k = -n; k <= n; ++k) dp[at(k)] = c(k) * full[at(k - 1)] - c(k + 1) * full[at(k + 1)];
      for (int m = 0; m <= n; ++m)
        out[N * idx(m, n) + i] = c(m) * dp[at(m - 1)] - c(m + 1) * dp[at(m + 1)];
    }
}

static void fill_fourier_d2(int p0, int p1, double *M) {
  const double s2p = std::sqrt(2 * kPI);
  for (int j = 0; j < 2 * p1; ++j) {
    M[j] = 0.0;
    for (int k = 1; k < p0; ++k) {
      M[(2 * k - 1) * 2 * p1 + j] = -s2p * k * k * std::cos(j * k * kPI / p1);
      M[(2 * k) * 2 * p1 + j] = -s2p * k * k * std::sin(j * k * kPI / p1);
    }
    M[(2 * p0 - 1) * 2 * p1 + j] = -s2p * p0 * p0 * std::cos(j * p0 * kPI / p1);
  }
}

static void fill_fourier(int p0, int p1, double *M) {  // src/SphericalHarmonics.cc:606-627
  const double s2p = std::sqrt(2 * kPI);
  for (int j = 0; j < 2 * p1; ++j) {
    M[0 * 2 * p1 + j] = s2p;
    for (int k = 1; k < p0; ++k) {
      M[(2 * k - 1) * 2 * p1 + j] = s2p * std::cos(j * k * kPI / p1);
      M[(2 * k) * 2 * p1 + j] = s2p * std::sin(j * k * kPI / p1);
    }
    M[(2 * p0 - 1) * 2 * p1 + j] = s2p * std::cos(j * p0 * kPI / p1);
  }
}

static void fill_fourier_grad(int p0, int p1, double *M) {  // :655-676
  const double s2p = std::sqrt(2 * kPI);
  for (int j = 0; j < 2 * p1; ++j) {
    M[j] = 0.0;
    for (int k = 1; k < p0; ++k) {
      M[(2 * k - 1) * 2 * p1 + j] = -s2p * k * std::sin(j * k * kPI / p1);
      M[(2 * k) * 2 * p1 + j] = s2p * k * std::cos(j * k * kPI / p1);
    }
    M[(2 * p0 - 1) * 2 * p1 + j] = -s2p * p0 * std::sin(j * p0 * kPI / p1);
  }
}

static void fill_fourier_inv(int p0, int p1, double *M) {  // :629-653
  const double c = 1.0 / std::sqrt(2 * kPI) / p0;
  const int cols = 2 * p1;
  for (int i = 0; i < 2 * p0 * cols; ++i) M[i] = 0.0;
  const int q = p1 > p0 ? p0 : p1;
  for (int j = 0; j < 2 * p0; ++j) {
    M[j * cols + 0] = c * 0.5;
    for (int k = 1; k < q; ++k) {
      M[j * cols + 2 * k - 1] = c * std::cos(j * k * kPI / p0);
      M[j * cols + 2 * k] = c * std::sin(j * k * kPI / p0);
    }
    M[j * cols + 2 * q - 1] = c * std::cos(j * q * kPI / p0);
    if (q == p0) M[j * cols + 2 * q - 1] *= 0.5;
  }
}

void build_sh_mats(int p0, int p1, ShMatsHost &h) {
  h.p0 = p0;
  h.p1 = p1;
  auto &D = h.data;
  D.clear();
  auto alloc = [&D](long n) {
    long o = (long)D.size();
    D.resize(D.size() + n, 0.0);
    return o;
  };
  h.Mf = alloc(4L * p0 * p1);
  h.Mdf = alloc(4L * p0 * p1);
  h.Mfinv = alloc(4L * p0 * p1);
  h.MfT = alloc(4L * p0 * p1);
  h.Md2f = alloc(4L * p0 * p1);
  fill_fourier_d2(p0, p1, &D[h.Md2f]);
  fill_fourier(p0, p1, &D[h.Mf]);
  fill_fourier_grad(p0, p1, &D[h.Mdf]);
  fill_fourier_inv(p0, p1, &D[h.Mfinv]);
  fill_fourier(p1, p0, &D[h.MfT]);

  std::vector<double> x1, w1, x0, w0;
  gauss_legendre(p1 + 1, x1, w1);
  gauss_legendre(p0 + 1, x0, w0);
  {  // MatLegendre(p0,p1), MatLegendreGrad(p0,p1): LegPoly(p0) at p1 nodes (:678-701, :735-757)
    std::vector<double> alp((long)(p0 + 1) * (p0 + 2) / 2 * (p1 + 1));
    std::vector<double> dlp(alp.size()), d2lp(alp.size());
    legpoly(x1.data(), p1 + 1, p0, alp.data());
    legpoly_deriv(x1.data(), p1 + 1, p0, dlp.data());
    legpoly_d2(x1.data(), p1 + 1, p0, d2lp.data());
    long src = 0;
    h.Ml.resize(p0 + 1);
    h.Mdl.resize(p0 + 1);
    h.Mdln.resize(p0 + 1);
    h.Md2l.resize(p0 + 1);
    for (int m = 0; m <= p0; ++m) {
      long n = (long)(p0 + 1 - m) * (p1 + 1);
      h.Ml[m] = alloc(n);
      h.Mdl[m] = alloc(n);
      h.Mdln[m] = alloc(n);
      h.Md2l[m] = alloc(n);
      for (long i = 0; i < n; ++i) {
        D[h.Ml[m] + i] = alp[src + i];
        D[h.Mdl[m] + i] = dlp[src + i];
        D[h.Mdln[m] + i] = -dlp[src + i];
        D[h.Md2l[m] + i] = d2lp[src + i];
      }
      src += n;
    }
  }
  {  // MatLegendreInv(p0,p1): LegPoly(p1) at p0 nodes, transposed, times w 2pi (:703-733)
    std::vector<double> alp((long)(p1 + 1) * (p1 + 2) / 2 * (p0 + 1));
    legpoly(x0.data(), p0 + 1, p1, alp.data());
    long src = 0;
    h.Mlinv.resize(p1 + 1);
    for (int m = 0; m <= p1; ++m) {
      int rows = p1 + 1 - m;  // degrees
      h.Mlinv[m] = alloc((long)(p0 + 1) * rows);
      for (int j = 0; j < rows; ++j)
        for (int k = 0; k <= p0; ++k)
          D[h.Mlinv[m] + (long)k * rows + j] = alp[src + (long)j * (p0 + 1) + k] * w0[k] * 2 * kPI;
      src += (long)rows * (p0 + 1);
    }
  }
  {  // MatLegendre(p1,p0): LegPoly(p1) at p0 nodes (read transposed)
    std::vector<double> alp((long)(p1 + 1) * (p1 + 2) / 2 * (p0 + 1));
    legpoly(x0.data(), p0 + 1, p1, alp.data());
    long src = 0;
    h.MlT.resize(p1 + 1);
    for (int m = 0; m <= p1; ++m) {
      long n = (long)(p1 + 1 - m) * (p0 + 1);
      h.MlT[m] = alloc(n);
      for (long i = 0; i < n; ++i) D[h.MlT[m] + i] = alp[src + i];
      src += n;
    }
  }
}

// Plain FP64 Grid2SHC (p0 -> p1 with the matrices of m = (p0, p1)).
void host_grid2shc(const ShMatsHost &m, const double *X, long N, double *S) {
  const int p0 = m.p0, p1 = m.p1;
  const int nc1 = sh_ncoef(p1);
  std::vector<double> B((long)2 * p1 * (p0 + 1));
  const double *Mf = &m.data[m.Mfinv];
  for (long f = 0; f < N; ++f) {
    const double *x = X + (long)f * sh_ngrid(p0);
    for (int lat = 0; lat <= p0; ++lat)
      for (int c = 0; c < 2 * p1; ++c) {
        double s = 0;
        for (int k = 0; k < 2 * p0; ++k) s += x[lat * 2 * p0 + k] * Mf[k * 2 * p1 + c];
        B[(long)c * (p0 + 1) + lat] = s;
      }
    for (int j = 0; j < 2 * p1; ++j) {
      int mm = (j + 1) / 2, len = sh_blocklen(p1, j), off = sh_blockoff(p1, j);
      const double *L = &m.data[m.Mlinv[mm]];
      for (int n = 0; n < len; ++n) {
        double s = 0;
        for (int k = 0; k <= p0; ++k) s += B[(long)j * (p0 + 1) + k] * L[(long)k * len + n];
        S[(long)f * nc1 + off + n] = s;
      }
    }
  }
}

void build_rot_mats(int p0, RotMatsHost &h) {  // src/SphericalHarmonics.cc:759-860
  h.p = p0;
  const int Ncoef = sh_ncoef(p0), Ngrid = sh_ngrid(p0), Naleg = (p0 + 1) * (p0 + 2) / 2;
  const double s2p = std::sqrt(2 * kPI);
  // coeff_perm: degree n holds coefficient indices perm_n[i], i over blocks j
  std::vector<std::vector<int>> cperm(p0 + 1);
  for (int n = 0; n <= p0; ++n) cperm[n].assign(std::min(2 * n + 1, 2 * p0), 0);
  {
    int itr = 0;
    for (int i = 0; i < 2 * p0; ++i) {
      int m = (i + 1) / 2;
      for (int n = m; n <= p0; ++n) cperm[n][i] = itr++;
    }
  }
  h.perm.clear();
  h.doff.assign(p0 + 2, 0);
  h.len.assign(p0 + 1, 0);
  for (int n = 0; n <= p0; ++n) {
    h.doff[n] = (int)h.perm.size();
    h.len[n] = (int)cperm[n].size();
    for (int v : cperm[n]) h.perm.push_back(v);
  }
  h.doff[p0 + 1] = (int)h.perm.size();

  std::vector<double> x, w;
  gauss_legendre(p0 + 1, x, w);
  std::vector<double> C0(3L * Ngrid);
  for (int i = 0; i <= p0; ++i)
    for (int j = 0; j < 2 * p0; ++j) {
      double s = std::sqrt(1 - x[i] * x[i]);
      C0[0L * Ngrid + i * 2 * p0 + j] = x[i];
      C0[1L * Ngrid + i * 2 * p0 + j] = s * std::sin(kPI * j / p0);
      C0[2L * Ngrid + i * 2 * p0 + j] = s * std::cos(kPI * j / p0);
    }
  ShMatsHost mats;
  build_sh_mats(p0, p0, mats);
  h.off.assign((p0 + 1) * (p0 + 1), 0);
  h.data.clear();
  std::vector<double> C1(3L * Ngrid), Mleg((long)Naleg * Ngrid), theta(Ngrid);
  std::vector<double> M((long)Ncoef * Ngrid), C2C((long)Ncoef * Ncoef);
  for (int l = 0; l <= p0; ++l) {
    double c = -x[l], sn = -std::sqrt(1.0 - x[l] * x[l]);
    for (int k = 0; k < Ngrid; ++k) {
      double a = C0[k], b = C0[Ngrid + k], d = C0[2L * Ngrid + k];
      C1[k] = c * a - sn * d;
      C1[Ngrid + k] = b;
      C1[2L * Ngrid + k] = sn * a + c * d;
    }
    legpoly(C1.data(), Ngrid, p0, Mleg.data());
    for (int k = 0; k < Ngrid; ++k) theta[k] = std::atan2(C1[Ngrid + k], C1[2L * Ngrid + k]);
    long o0 = 0, o1 = 0;
    for (int i = 0; i <= p0; ++i) {
      int len = p0 + 1 - i;
      for (int j = 0; j < len; ++j)
        for (int k = 0; k < Ngrid; ++k)
          M[(o1 + j) * Ngrid + k] = s2p * Mleg[(o0 + j) * Ngrid + k] * std::cos(i * theta[k]);
      o1 += len;
      if (i != 0 && i != p0) {
        for (int j = 0; j < len; ++j)
          for (int k = 0; k < Ngrid; ++k)
            M[(o1 + j) * Ngrid + k] = s2p * Mleg[(o0 + j) * Ngrid + k] * std::sin(i * theta[k]);
        o1 += len;
      }
      o0 += len;
    }
    host_grid2shc(mats, M.data(), Ncoef, C2C.data());
    for (int n = 0; n <= p0; ++n) {
      const auto &pn = cperm[n];
      const int L = (int)pn.size();
      h.off[l * (p0 + 1) + n] = (long)h.data.size();
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) h.data.push_back(C2C[(long)pn[i] * Ncoef + pn[j]]);
    }
  }
}

void legendre_weights(int p, std::vector<double> &qw) {  // :557-571
  std::vector<double> x;
  gauss_legendre(p + 1, x, qw);
  for (int i = 0; i <= p; ++i) qw[i] *= kPI / p / std::sqrt(1 - x[i] * x[i]);
}

void singular_weights(int p1, std::vector<double> &sw) {  // :573-604
  std::vector<double> x, w;
  gauss_legendre(p1 + 1, x, w);
  std::vector<double> alp0((p1 + 1) * (p1 + 2) / 2);
  double one = 1.0;
  legpoly(&one, 1, p1, alp0.data());
  std::vector<double> alp((long)(p1 + 1) * (p1 + 1) * (p1 + 2) / 2);
  legpoly(x.data(), p1 + 1, p1, alp.data());
  std::vector<double> Yf(p1 + 1, 0.0);
  for (int j = 0; j <= p1; ++j)
    for (int i = 0; i <= p1; ++i) Yf[i] += 4 * kPI / (2 * j + 1) * alp0[j] * alp[(long)j * (p1 + 1) + i];
  sw.resize(p1 + 1);
  for (int i = 0; i <= p1; ++i) sw[i] = (w[i] * kPI / p1) * Yf[i] / std::cos(std::acos(x[i]) / 2);
}

// ------------------------------------------------------------ device caches
namespace {
std::mutex g_mu;
std::map<std::pair<int, int>, std::unique_ptr<ShMatsHost>> g_host;
std::map<std::pair<int, int>, std::unique_ptr<ShMatsDev>> g_dev;  // key (device, p0*1024+p1)
std::map<std::pair<int, int>, std::unique_ptr<RotMatsHost>> g_rhost;
std::map<std::pair<int, int>, std::unique_ptr<RotMatsDev>> g_rdev;

template <typename T>
T *upload(const std::vector<T> &v) {
  T *d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (v.empty() ? 1 : v.size())) != hipSuccess) return nullptr;
  if (!v.empty() && hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

std::vector<float> to_f32(const std::vector<double> &v) {
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
  return f;
}
}  // namespace

const ShMatsDev *sh_mats(int p0, int p1) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, p0 * 1024 + p1);
  auto it = g_dev.find(key);
  if (it != g_dev.end()) return it->second.get();
  auto hk = std::make_pair(p0, p1);
  if (!g_host.count(hk)) {
    auto h = std::make_unique<ShMatsHost>();
    build_sh_mats(p0, p1, *h);
    g_host[hk] = std::move(h);
  }
  auto d = std::make_unique<ShMatsDev>();
  d->p0 = p0;
  d->p1 = p1;
  d->h = g_host[hk].get();
  d->d64 = upload(d->h->data);
  d->d32 = upload(to_f32(d->h->data));
  if (!d->d64 || !d->d32) return nullptr;
  ShMatsDev *r = d.get();
  g_dev[key] = std::move(d);
  return r;
}

const RotMatsDev *rot_mats(int p) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, p);
  auto it = g_rdev.find(key);
  if (it != g_rdev.end()) return it->second.get();
  auto hk = std::make_pair(p, 0);
  if (!g_rhost.count(hk)) {
    auto h = std::make_unique<RotMatsHost>();
    build_rot_mats(p, *h);
    g_rhost[hk] = std::move(h);
  }
  const RotMatsHost *h = g_rhost[hk].get();
  const int Ncoef = sh_ncoef(p);
  std::vector<int> iperm(Ncoef), partner(Ncoef), order(Ncoef), is_sin(Ncoef);
  for (int i = 0; i < Ncoef; ++i) iperm[h->perm[i]] = i;
  for (int j = 0; j < 2 * p; ++j) {
    int m = (j + 1) / 2, len = sh_blocklen(p, j), off = sh_blockoff(p, j);
    for (int n = 0; n < len; ++n) {
      int c = off + n;
      order[c] = m;
      bool sn = (m != 0 && m != p && j == 2 * m);
      is_sin[c] = sn ? 1 : 0;
      if (m == 0 || m == p) partner[c] = c;
      else partner[c] = sn ? c - len : c + len;
    }
  }
  std::vector<double> phase(2L * p * (p + 1));
  for (int j = 0; j < p; ++j)
    for (int k = 0; k <= p; ++k) {
      const double a = (double)k * j * kPI / p;
      phase[2 * (j * (p + 1) + k)] = std::cos(a);
      phase[2 * (j * (p + 1) + k) + 1] = std::sin(a);
    }
  std::vector<int> scatter(4L * Ncoef);
  for (int c = 0; c < Ncoef; ++c) {
    int *e = &scatter[4L * c];
    const int k = order[c];
    e[2] = k;
    if (k == 0 || k == p) {
      e[0] = iperm[c];
      e[1] = 0;
      e[3] = 0;
    } else {
      const bool sn = is_sin[c] != 0;
      e[0] = iperm[sn ? partner[c] : c];
      e[1] = iperm[sn ? c : partner[c]];
      e[3] = sn ? 2 : 1;
    }
  }
  auto d = std::make_unique<RotMatsDev>();
  d->h = h;
  d->phase = upload(phase);
  d->scatter = upload(scatter);
  d->d64 = upload(h->data);
  d->d32 = upload(to_f32(h->data));
  d->perm = upload(h->perm);
  d->iperm = upload(iperm);
  d->partner = upload(partner);
  d->order = upload(order);
  d->is_sin = upload(is_sin);
  if (!d->d64 || !d->d32 || !d->perm || !d->iperm || !d->partner || !d->order || !d->is_sin ||
      !d->phase || !d->scatter)
    return nullptr;
  RotMatsDev *r = d.get();
  g_rdev[key] = std::move(d);
  return r;
}

}  // namespace v3d

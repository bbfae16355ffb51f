// sht.hip -- SphericalHarmonics transforms as grouped GEMMs.
//
// SHC2Grid (src/SphericalHarmonics.cc:3-149):
//   Legendre: for every coefficient block j (order m=(j+1)/2, cos or sin)
//     B1[j][f][lat] = sum_n S[f][off_j + n] * Ml[m][n][lat]
//   Fourier : X[f][lat][lon] = sum_j B1[j][f][lat] * Mf[j][lon]
// Grid2SHC (:151-235) and SHC2GridTranspose (:237-322) run the same two stages
// in the opposite order with the inverse / transposed matrices.  The block
// gather of the coefficient layout and the transposes between the stages are
// operand strides of the GEMM descriptors; nothing is re-laid out in memory.
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <cmath>

#include "runtime.hpp"
#include "sht.hpp"

namespace v3d {

namespace {

// Synthesis stages carry a derivative variant v in {0, 1, 2} (value, first,
// second derivative) folded into the kind as kind + 16 * v.
enum PlanKind { kS2GLeg, kS2GFour, kG2SFour, kG2SLeg, kTFour, kTLeg };

struct Plan {
  GemmDesc *d = nullptr;
  int n = 0, max_m = 0, max_n = 0, tile = 0;
};

std::mutex g_mu;
std::map<std::tuple<int, int, int, int, long>, Plan> g_plans;  // (device, kind, p0, p1, N)

Plan make_plan(int kindv, int p0, int p1, long N, const ShMatsHost &h) {
  const int kind = kindv % 16, var = kindv / 16;
  std::vector<GemmDesc> v;
  const int nc0 = sh_ncoef(p0), nc1 = sh_ncoef(p1);
  auto push = [&v](int M, int Nn, int K, long ao, long ars, long acs, long bo, long brs,
                   long bcs, long co, long crs, long ccs) {
    GemmDesc d{};
    d.M = M; d.N = Nn; d.K = K;
    d.a_off = ao; d.a_rs = ars; d.a_cs = acs; d.a_bs = 0;
    d.b_off = bo; d.b_rs = brs; d.b_cs = bcs; d.b_bs = 0;
    d.c_off = co; d.c_rs = crs; d.c_cs = ccs; d.c_bs = 0;
    v.push_back(d);
  };
  switch (kind) {
    case kS2GLeg:
      for (int j = 0; j < 2 * p0; ++j) {
        const int m = (j + 1) / 2;
        const long bo = var == 0 ? h.Ml[m] : (var == 1 ? h.Mdl[m] : h.Md2l[m]);
        push((int)N, p1 + 1, sh_blocklen(p0, j), sh_blockoff(p0, j), nc0, 1, bo, p1 + 1, 1,
             (long)j * N * (p1 + 1), p1 + 1, 1);
      }
      break;
    case kS2GFour:
      push((int)(N * (p1 + 1)), 2 * p1, 2 * p0, 0, 1, N * (p1 + 1),
           var == 0 ? h.Mf : (var == 1 ? h.Mdf : h.Md2f), 2 * p1, 1, 0, 2 * p1, 1);
      break;
    case kG2SFour:
      push((int)(N * (p0 + 1)), 2 * p1, 2 * p0, 0, 2 * p0, 1, h.Mfinv, 2 * p1, 1, 0, 1,
           N * (p0 + 1));
      break;
    case kTFour:  // MatFourier(p1,p0) stored [2p1][2p0], read transposed
      push((int)(N * (p0 + 1)), 2 * p1, 2 * p0, 0, 2 * p0, 1, h.MfT, 1, 2 * p0, 0, 1,
           N * (p0 + 1));
      break;
    case kG2SLeg:
    case kTLeg:
      for (int j = 0; j < 2 * p1; ++j) {
        const int m = (j + 1) / 2;
        const int len = sh_blocklen(p1, j);
        if (kind == kG2SLeg)
          push((int)N, len, p0 + 1, (long)j * N * (p0 + 1), p0 + 1, 1, h.Mlinv[m], len, 1,
               sh_blockoff(p1, j), nc1, 1);
        else  // MatLegendre(p1,p0)[m] stored [(p1+1-m)][(p0+1)], read transposed
          push((int)N, len, p0 + 1, (long)j * N * (p0 + 1), p0 + 1, 1, h.MlT[m], 1, p0 + 1,
               sh_blockoff(p1, j), nc1, 1);
      }
      break;
  }
  Plan p;
  p.n = (int)v.size();
  for (auto &d : v) {
    p.max_m = std::max(p.max_m, d.M);
    p.max_n = std::max(p.max_n, d.N);
  }
  p.tile = gemm_pick_tile(v.data(), p.n);
  p.d = upload_descs(v.data(), p.n);
  return p;
}

const Plan *get_plan(int kind, int p0, int p1, long N, const ShMatsHost &h) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, kind, p0, p1, N);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return &it->second;
  Plan p = make_plan(kind, p0, p1, N, h);
  if (!p.d) return nullptr;
  return &(g_plans[key] = p);
}

template <typename T>
const T *mats_ptr(const ShMatsDev *m);
template <>
const double *mats_ptr<double>(const ShMatsDev *m) { return m->d64; }
template <>
const float *mats_ptr<float>(const ShMatsDev *m) { return m->d32; }

template <typename T>
int run(int kind, int p0, int p1, long N, const ShMatsDev *m, const T *A, T *C,
        hipStream_t st, T alpha = T(1)) {
  const Plan *pl = get_plan(kind, p0, p1, N, *m->h);
  if (!pl) return VES3D_MEMORY_ERROR;
  return gemm_grouped<T>(pl->d, pl->n, 1, pl->max_m, pl->max_n, A, mats_ptr<T>(m), C,
                         alpha, T(0), st, pl->tile);
}

bool valid_order(int p) { return p >= 1 && p <= 255; }

// plans of sht_synth_multi: the Legendre and the Fourier stage of every
// variant, keyed by (device, p0, p1, N, variants, x_stride)
struct MultiPlan {
  Plan leg, four;
};
std::map<std::tuple<int, int, int, long, int, long>, MultiPlan> g_mplans;

Plan finish_plan(std::vector<GemmDesc> &v) {
  Plan p;
  p.n = (int)v.size();
  for (auto &d : v) {
    p.max_m = std::max(p.max_m, d.M);
    p.max_n = std::max(p.max_n, d.N);
  }
  p.tile = gemm_pick_tile(v.data(), p.n);
  p.d = upload_descs(v.data(), p.n);
  return p;
}

const MultiPlan *get_multi_plan(int p0, int p1, long N, int nvar, const int *dleg,
                                const int *dfour, const int *neg, long x_stride,
                                const ShMatsHost &h) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  int code = nvar;
  for (int v = 0; v < nvar; ++v)
    code = code * 32 + dleg[v] * 8 + dfour[v] * 2 + (neg[v] ? 1 : 0);
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, p0, p1, N, code, x_stride);
  auto it = g_mplans.find(key);
  if (it != g_mplans.end()) return &it->second;
  const int nc0 = sh_ncoef(p0);
  const long slab = shc2grid_work(N, p0, p1);
  std::vector<GemmDesc> leg, four;
  auto desc = [](int M, int Nn, int K, long ao, long ars, long acs, long bo, long brs, long bcs,
                 long co, long crs, long ccs) {
    GemmDesc d{};
    d.M = M; d.N = Nn; d.K = K;
    d.a_off = ao; d.a_rs = ars; d.a_cs = acs;
    d.b_off = bo; d.b_rs = brs; d.b_cs = bcs;
    d.c_off = co; d.c_rs = crs; d.c_cs = ccs;
    return d;
  };
  for (int v = 0; v < nvar; ++v) {
    for (int j = 0; j < 2 * p0; ++j) {
      const int m = (j + 1) / 2;
      const long bo = dleg[v] == 0   ? h.Ml[m]
                      : dleg[v] == 1 ? (neg[v] ? h.Mdln[m] : h.Mdl[m])
                                     : h.Md2l[m];
      leg.push_back(desc((int)N, p1 + 1, sh_blocklen(p0, j), sh_blockoff(p0, j), nc0, 1, bo,
                         p1 + 1, 1, v * slab + (long)j * N * (p1 + 1), p1 + 1, 1));
    }
    four.push_back(desc((int)(N * (p1 + 1)), 2 * p1, 2 * p0, v * slab, 1, N * (p1 + 1),
                        dfour[v] == 0 ? h.Mf : (dfour[v] == 1 ? h.Mdf : h.Md2f), 2 * p1, 1,
                        v * x_stride, 2 * p1, 1));
  }
  MultiPlan mp;
  mp.leg = finish_plan(leg);
  mp.four = finish_plan(four);
  if (!mp.leg.d || !mp.four.d) return nullptr;
  return &(g_mplans[key] = mp);
}

}  // namespace

template <typename T>
int sht_synth_multi(const T *S, long N, int p0, int p1, int nvar, const int *dleg,
                    const int *dfour, const int *neg, T *X, long x_stride, T *work,
                    hipStream_t st) {
  if (!valid_order(p0) || !valid_order(p1) || N < 0 || nvar < 1 || nvar > 4 ||
      x_stride < N * sh_ngrid(p1))
    return VES3D_INVALID_PARAMETER;
  for (int v = 0; v < nvar; ++v)
    if (dleg[v] < 0 || dleg[v] > 2 || dfour[v] < 0 || dfour[v] > 2 || (neg[v] && dleg[v] != 1))
      return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return VES3D_MEMORY_ERROR;
  const MultiPlan *mp = get_multi_plan(p0, p1, N, nvar, dleg, dfour, neg, x_stride, *m->h);
  if (!mp) return VES3D_MEMORY_ERROR;
  V3D_CHECK(gemm_grouped<T>(mp->leg.d, mp->leg.n, 1, mp->leg.max_m, mp->leg.max_n, S,
                            mats_ptr<T>(m), work, T(1), T(0), st, mp->leg.tile));
  return gemm_grouped<T>(mp->four.d, mp->four.n, 1, mp->four.max_m, mp->four.max_n, work,
                         mats_ptr<T>(m), X, T(1), T(0), st, mp->four.tile);
}

template <typename T>
int sht_shc2grid(const T *S, long N, int p0, int p1, T *X, T *Xt, T *Xp, T *work,
                 hipStream_t st) {
  if (!valid_order(p0) || !valid_order(p1) || N < 0) return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return VES3D_MEMORY_ERROR;
  V3D_CHECK(run<T>(kS2GLeg, p0, p1, N, m, S, work, st));
  V3D_CHECK(run<T>(kS2GFour, p0, p1, N, m, work, X, st));
  if (Xt) V3D_CHECK(run<T>(kS2GFour + 16, p0, p1, N, m, work, Xt, st));
  if (Xp) {
    V3D_CHECK(run<T>(kS2GLeg + 16, p0, p1, N, m, S, work, st));
    V3D_CHECK(run<T>(kS2GFour, p0, p1, N, m, work, Xp, st));
  }
  return VES3D_SUCCESS;
}

template <typename T>
int sht_synth(const T *S, long N, int p0, int p1, int dleg, int dfour, T *X, T *work,
              hipStream_t st, T alpha) {
  if (!valid_order(p0) || !valid_order(p1) || N < 0 || dleg < 0 || dleg > 2 || dfour < 0 ||
      dfour > 2)
    return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return VES3D_MEMORY_ERROR;
  V3D_CHECK(run<T>(kS2GLeg + 16 * dleg, p0, p1, N, m, S, work, st, alpha));
  return run<T>(kS2GFour + 16 * dfour, p0, p1, N, m, work, X, st);
}

template <typename T>
int sht_grid2shc(const T *X, long N, int p0, int p1, T *S, T *work, hipStream_t st) {
  if (!valid_order(p0) || !valid_order(p1) || N < 0) return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return VES3D_MEMORY_ERROR;
  V3D_CHECK(run<T>(kG2SFour, p0, p1, N, m, X, work, st));
  return run<T>(kG2SLeg, p0, p1, N, m, work, S, st);
}

template <typename T>
int sht_shc2grid_transpose(const T *X, long N, int p0, int p1, T *S, T *work,
                           hipStream_t st) {
  if (!valid_order(p0) || !valid_order(p1) || N < 0) return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return VES3D_MEMORY_ERROR;
  V3D_CHECK(run<T>(kTFour, p0, p1, N, m, X, work, st));
  return run<T>(kTLeg, p0, p1, N, m, work, S, st);
}

template <typename T>
const T *sht_mft(int p0, int p1) {
  if (!valid_order(p0) || !valid_order(p1)) return nullptr;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return nullptr;
  return mats_ptr<T>(m) + m->h->MfT;
}

template <typename T>
const T *sht_mlt(int p0, int p1) {
  if (!valid_order(p0) || !valid_order(p1)) return nullptr;
  const ShMatsDev *m = sh_mats(p0, p1);
  if (!m) return nullptr;
  return mats_ptr<T>(m) + m->h->MlT[0];
}

// ---------------------------------------------------------------- poles
namespace {
std::map<std::pair<int, int>, double *> g_qp;  // (device, p) -> [2][p+1]

const double *pole_table(int p) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, p);
  auto it = g_qp.find(key);
  if (it != g_qp.end()) return it->second;
  std::vector<double> h(2 * (p + 1)), alp((p + 1) * (p + 2) / 2);
  const double s2p = std::sqrt(2 * kPi);
  for (int t = 0; t < 2; ++t) {
    double x = t == 0 ? -1.0 : 1.0;
    legpoly(&x, 1, p, alp.data());
    for (int j = 0; j <= p; ++j) h[t * (p + 1) + j] = alp[j] * s2p;
  }
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  g_qp[key] = d;
  return d;
}

template <typename T>
__global__ void k_shc2pole(const T *__restrict__ S, long N, int p0, int ncoef,
                           const double *__restrict__ qp, T *__restrict__ P) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * N) return;
  const long f = i >> 1;
  const int t = (int)(i & 1);
  double s = 0;
  for (int j = 0; j <= p0; ++j) s += (double)S[f * ncoef + j] * qp[t * (p0 + 1) + j];
  P[i] = (T)s;
}
}  // namespace

template <typename T>
int sht_shc2pole(const T *S, long N, int p0, T *P, hipStream_t st) {
  if (!valid_order(p0) || N < 0) return VES3D_INVALID_PARAMETER;
  if (N == 0) return VES3D_SUCCESS;
  const double *qp = pole_table(p0);
  if (!qp) return VES3D_MEMORY_ERROR;
  k_shc2pole<T><<<cdiv(2 * N, 256), 256, 0, st>>>(S, N, p0, sh_ncoef(p0), qp, P);
  return launch_status();
}

#define V3D_INST(T)                                                                       \
  template int sht_shc2grid<T>(const T *, long, int, int, T *, T *, T *, T *, hipStream_t); \
  template int sht_grid2shc<T>(const T *, long, int, int, T *, T *, hipStream_t);          \
  template int sht_shc2grid_transpose<T>(const T *, long, int, int, T *, T *, hipStream_t); \
  template int sht_shc2pole<T>(const T *, long, int, T *, hipStream_t);                 \
  template int sht_synth<T>(const T *, long, int, int, int, int, T *, T *, hipStream_t, T); \
  template int sht_synth_multi<T>(const T *, long, int, int, int, const int *, const int *,    \
                                  const int *, T *, long, T *, hipStream_t);                   \
  template const T *sht_mft<T>(int, int);                                                  \
  template const T *sht_mlt<T>(int, int);
V3D_INST(double)
V3D_INST(float)
#undef V3D_INST

}  // namespace v3d

// ---------------------------------------------------------------- C ABI
using namespace v3d;

namespace {
template <typename T>
int abi_shc2grid(const T *S, size_t N, int p0, int p1, T *X, T *Xt, T *Xp, void *stream) {
  if (!S || !X) return VES3D_INVALID_PARAMETER;
  T *w = (T *)workspace(kWsSht0, sizeof(T) * shc2grid_work((long)N, p0, p1));
  if (!w) return VES3D_MEMORY_ERROR;
  return sht_shc2grid<T>(S, (long)N, p0, p1, X, Xt, Xp, w, (hipStream_t)stream);
}
template <typename T>
int abi_grid2shc(const T *X, size_t N, int p0, int p1, T *S, void *stream) {
  if (!S || !X) return VES3D_INVALID_PARAMETER;
  T *w = (T *)workspace(kWsSht0, sizeof(T) * grid2shc_work((long)N, p0, p1));
  if (!w) return VES3D_MEMORY_ERROR;
  return sht_grid2shc<T>(X, (long)N, p0, p1, S, w, (hipStream_t)stream);
}
template <typename T>
int abi_transpose(const T *X, size_t N, int p0, int p1, T *S, void *stream) {
  if (!S || !X) return VES3D_INVALID_PARAMETER;
  T *w = (T *)workspace(kWsSht0, sizeof(T) * grid2shc_work((long)N, p0, p1));
  if (!w) return VES3D_MEMORY_ERROR;
  return sht_shc2grid_transpose<T>(X, (long)N, p0, p1, S, w, (hipStream_t)stream);
}
}  // namespace

extern "C" {
int ves3d_hip_shc2grid_f64(const double *S, size_t N, int p0, int p1, double *X,
                           double *X_theta, double *X_phi, ves3d_stream_t stream) {
  return abi_shc2grid<double>(S, N, p0, p1, X, X_theta, X_phi, stream);
}
int ves3d_hip_shc2grid_f32(const float *S, size_t N, int p0, int p1, float *X,
                           float *X_theta, float *X_phi, ves3d_stream_t stream) {
  return abi_shc2grid<float>(S, N, p0, p1, X, X_theta, X_phi, stream);
}
int ves3d_hip_grid2shc_f64(const double *X, size_t N, int p0, int p1, double *S,
                           ves3d_stream_t stream) {
  return abi_grid2shc<double>(X, N, p0, p1, S, stream);
}
int ves3d_hip_grid2shc_f32(const float *X, size_t N, int p0, int p1, float *S,
                           ves3d_stream_t stream) {
  return abi_grid2shc<float>(X, N, p0, p1, S, stream);
}
int ves3d_hip_shc2grid_transpose_f64(const double *X, size_t N, int p0, int p1,
                                     double *S, ves3d_stream_t stream) {
  return abi_transpose<double>(X, N, p0, p1, S, stream);
}
int ves3d_hip_shc2grid_transpose_f32(const float *X, size_t N, int p0, int p1,
                                     float *S, ves3d_stream_t stream) {
  return abi_transpose<float>(X, N, p0, p1, S, stream);
}
int ves3d_hip_shc2pole_f64(const double *S, size_t N, int p0, double *P,
                           ves3d_stream_t stream) {
  if (!S || !P) return VES3D_INVALID_PARAMETER;
  return sht_shc2pole<double>(S, (long)N, p0, P, (hipStream_t)stream);
}
int ves3d_hip_shc2pole_f32(const float *S, size_t N, int p0, float *P,
                           ves3d_stream_t stream) {
  if (!S || !P) return VES3D_INVALID_PARAMETER;
  return sht_shc2pole<float>(S, (long)N, p0, P, (hipStream_t)stream);
}
}

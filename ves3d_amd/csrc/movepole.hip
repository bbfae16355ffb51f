// movepole.hip -- the old pole-moving route of the singular quadrature:
// MovePole (src/MovePole.cc) with its Direct, DirectEagerEval and ViaSpHarm
// schemes and Device<>::AggregateRotation (src/DeviceCPU.cc:644-681,
// src/DeviceGPU.cc:312-339) on gfx950, FP64.
//
// Grid fields are [nsub][np] (np = 2p(p+1), point lat*2p + lon).
//   Direct       out = X_j R_i^T: X_j is X with its longitude index advanced
//                by j inside every latitude ring (the CircShift of the
//                reference's matrix moved onto the operand), then one GEMM.
//   Aggregate    all 2p longitudes of latitude i in ONE GEMM: the 2p shifted
//                copies of X are stacked as rows (2p nsub x np) against R_i^T.
//   ViaSpHarm    SHTrans forward + collectSameOrder once per operand; per
//                (i, j): longitude rotation of the (cos, sin) pairs, one
//                grouped GEMM over the p+1 per-degree blocks,
//                collectSameFreq, SHTrans backward.
// The rotation matrices are the reference's precomputed files when given,
// otherwise built here from their definition (any p; the files exist for
// p = 6 only): R_i = E(R_y(theta_i - pi) grid) Grid2SHC and the degree
// blocks of Grid2SHC R_i E(grid).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "gemm.hpp"
#include "sh_tables.hpp"

extern "C" int ves3d_hip_shtrans_forward(int, size_t, const double *, double *, void *);
extern "C" int ves3d_hip_shtrans_backward(int, size_t, const double *, int, int, double *, void *);

namespace v3d {
namespace {

inline unsigned nb(long n) {
  long b = (n + 255) / 256;
  return (unsigned)std::max(1L, std::min(b, 65536L));
}
#define MGL(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)(n); i += (long)gridDim.x * blockDim.x)

inline int deg_len(int p, int n) { return 2 * n + 1 - n / p; }

// ---------------------------------------------------------------- host tables
// Real SH basis of the SphericalHarmonics layout at (x, phi): [npts][p(p+2)].
void sh_basis(int p, const std::vector<double> &x, const std::vector<double> &phi,
              std::vector<double> &B) {
  const long n = (long)x.size();
  const int nc = sh_ncoef(p), nal = (p + 1) * (p + 2) / 2;
  std::vector<double> alp((long)nal * n);
  legpoly(x.data(), n, p, alp.data());
  std::vector<int> start(p + 1, 0);
  for (int m = 1; m <= p; ++m) start[m] = start[m - 1] + (p - m + 2);
  const double s2p = std::sqrt(2.0 * kPi);
  B.assign((size_t)n * nc, 0.0);
  for (int j = 0; j < 2 * p; ++j) {
    const int m = (j + 1) / 2, off = sh_blockoff(p, j), len = sh_blocklen(p, j);
    for (long q = 0; q < n; ++q) {
      double F;
      if (j == 0)
        F = 1.0;
      else if (j == 2 * p - 1)
        F = std::cos(p * phi[q]);
      else if (j % 2 == 1)
        F = std::cos(m * phi[q]);
      else
        F = std::sin(m * phi[q]);
      for (int k = 0; k < len; ++k) B[q * nc + off + k] = s2p * alp[(long)(start[m] + k) * n + q] * F;
    }
  }
}

// all_rot_mats [p+1][np][np], row = output point (matlab/GenerateDirectRotMat.m).
void build_direct_mats(int p, std::vector<double> &R) {
  const int np = sh_ngrid(p), nc = sh_ncoef(p);
  std::vector<double> x, w;
  gauss_legendre(p + 1, x, w);
  ShMatsHost mats;
  build_sh_mats(p, p, mats);
  std::vector<double> I((size_t)np * np, 0.0), G((size_t)np * nc);
  for (int k = 0; k < np; ++k) I[(size_t)k * np + k] = 1.0;
  host_grid2shc(mats, I.data(), np, G.data());     // G[k][c]: coefficients of the delta at k
  R.assign((size_t)(p + 1) * np * np, 0.0);
  std::vector<double> qx(np), qphi(np), B;
  for (int i = 0; i <= p; ++i) {
    const double a = std::acos(x[i]) - kPi, c = std::cos(a), s = std::sin(a);
    for (int lat = 0; lat <= p; ++lat)
      for (int lon = 0; lon < 2 * p; ++lon) {
        const int k = lat * 2 * p + lon;
        const double z = x[lat], rho = std::sqrt(1.0 - z * z), ph = kPi * lon / p;
        const double px = rho * std::cos(ph), py = rho * std::sin(ph);
        const double X = c * px + s * z, Y = py, Z = -s * px + c * z;
        qx[k] = std::max(-1.0, std::min(1.0, Z));
        qphi[k] = std::atan2(Y, X);
      }
    sh_basis(p, qx, qphi, B);
    double *Ri = &R[(size_t)i * np * np];
    for (int r = 0; r < np; ++r)
      for (int k = 0; k < np; ++k) {
        double acc = 0.0;
        for (int cc = 0; cc < nc; ++cc) acc += B[(size_t)r * nc + cc] * G[(size_t)k * nc + cc];
        Ri[(size_t)r * np + k] = acc;
      }
  }
}

// SpHarmRotMats: per latitude, per degree n a [len][len] block (row = output
// coefficient, SHTrans order within the degree: m = 0 cos, 1 cos, 1 sin, ...).
void build_spharm_mats(int p, const std::vector<double> &R, std::vector<double> &out) {
  const int np = sh_ngrid(p), nc = sh_ncoef(p);
  ShMatsHost mats;
  build_sh_mats(p, p, mats);
  std::vector<double> x, w;
  gauss_legendre(p + 1, x, w);
  std::vector<double> gx(np), gphi(np), E;
  for (int k = 0; k < np; ++k) {
    gx[k] = x[k / (2 * p)];
    gphi[k] = kPi * (k % (2 * p)) / p;
  }
  sh_basis(p, gx, gphi, E);                          // E[k][c]: synthesis at the grid
  out.clear();
  std::vector<double> rotf((size_t)nc * np), coef((size_t)nc * nc);
  for (int i = 0; i <= p; ++i) {
    const double *Ri = &R[(size_t)i * np * np];
    // rotf[c][r] = sum_k R_i[r][k] E[k][c]
    for (int c = 0; c < nc; ++c)
      for (int r = 0; r < np; ++r) {
        double acc = 0.0;
        for (int k = 0; k < np; ++k) acc += Ri[(size_t)r * np + k] * E[(size_t)k * nc + c];
        rotf[(size_t)c * np + r] = acc;
      }
    host_grid2shc(mats, rotf.data(), nc, coef.data());   // coef[in][out]
    for (int n = 0; n <= p; ++n) {
      const int L = deg_len(p, n);
      std::vector<int> idx(L);
      for (int jj = 0; jj < L; ++jj) idx[jj] = sh_blockoff(p, jj) + (n - (jj + 1) / 2);
      for (int ro = 0; ro < L; ++ro)
        for (int ci = 0; ci < L; ++ci) out.push_back(coef[(size_t)idx[ci] * nc + idx[ro]]);
    }
  }
}

// ------------------------------------------------------------------ kernels
// Y[j][s][lat*K + lon] = X[s][lat*K + (lon + j) mod K] for j in [j0, j0+nj).
__global__ void k_shift_lon(const double *__restrict__ X, long nsub, int p, int j0, int nj,
                            double *Y) {
  const long np = 2L * p * (p + 1), K = 2L * p;
  MGL(e, (long)nj * nsub * np) {
    const long k = e % np, s = (e / np) % nsub, j = j0 + e / (np * nsub);
    const long lat = k / K, lon = k % K;
    Y[e] = X[s * np + lat * K + (lon + j) % K];
  }
}

// Same-order <-> order-major coefficient layouts (SHTrans::collectSameOrder /
// collectSameFreq, src/SHTrans.cc:234-295).  Element e of the same-order
// array: degree n, function s, index jj in the degree; order-major offset of
// (block jj, function s, degree index n - m).
__device__ __forceinline__ long om_offset(int p, long ns, int jj, long s, int n) {
  long off = 0;
  for (int j = 0; j < jj; ++j) off += ns * (p + 1 - (j + 1) / 2);
  return off + s * (p + 1 - (jj + 1) / 2) + (n - (jj + 1) / 2);
}
__global__ void k_same_order(const double *__restrict__ in, int p, long ns, const int *dof,
                             int to_same, double *out) {
  const long tot = ns * (long)p * (p + 2);
  MGL(e, tot) {
    int n = 0;
    while (n < p && e >= ns * dof[n + 1]) ++n;
    const long r = e - ns * dof[n];
    const int L = 2 * n + 1 - n / p;
    const long s = r / L;
    const int jj = (int)(r % L);
    const long o = om_offset(p, ns, jj, s, n);
    if (to_same)
      out[e] = in[o];
    else
      out[o] = in[e];
  }
}

// alignMeridian (src/MovePole.cc:230-270): per order m the (cos, sin) pair
// (a, b) -> (c a + s b, -s a + c b), c = cos m lambda, s = sin m lambda.
__global__ void k_lon_rotate(const double *__restrict__ in, int p, long ns, const int *dof,
                             double lambda, double *out) {
  const long tot = ns * (long)p * (p + 2);
  MGL(e, tot) {
    int n = 0;
    while (n < p && e >= ns * dof[n + 1]) ++n;
    const long r = e - ns * dof[n];
    const int L = 2 * n + 1 - n / p;
    const int jj = (int)(r % L);
    const long base = e - jj;
    double v;
    if (jj == 0) {
      v = in[e];
    } else if (n == p && jj == 2 * p - 1) {
      v = cos(p * lambda) * in[e];
    } else {
      const int m = (jj + 1) / 2;
      const double c = cos(m * lambda), s = sin(m * lambda);
      const double a = in[base + 2 * m - 1], b = in[base + 2 * m];
      v = (jj % 2 == 1) ? c * a + s * b : -s * a + c * b;
    }
    out[e] = v;
  }
}

struct DevBuf {
  double *p = nullptr;
  size_t n = 0;
  int reserve(size_t count) {
    if (count <= n && p) return VES3D_SUCCESS;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, sizeof(double) * std::max<size_t>(count, 1)) != hipSuccess) {
      p = nullptr;
      return VES3D_MEMORY_ERROR;
    }
    n = std::max<size_t>(count, 1);
    return VES3D_SUCCESS;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Device copies of descriptor lists, uploaded once (synchronously) per
// distinct list and kept for the process lifetime.
const GemmDesc *cached_descs(const std::vector<GemmDesc> &h) {
  static std::mutex mu;
  static std::map<std::string, GemmDesc *> cache;
  const std::string key(reinterpret_cast<const char *>(h.data()), sizeof(GemmDesc) * h.size());
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  GemmDesc *d = upload_descs(h.data(), (int)h.size());
  if (d) cache[key] = d;
  return d;
}

// out[nrows][np] = X[nrows][np] R^T with R row-major [np][np] (rm = 1) or
// column-major (the reference's all_rot_mats blocks, rm = 0).
int rot_gemm(const double *X, long nrows, int np, const double *R, bool rm, double *out,
             hipStream_t st) {
  GemmDesc h{};
  h.M = (int)nrows;
  h.N = np;
  h.K = np;
  h.a_rs = np;
  h.a_cs = 1;
  // B(k, c) = R[c][k]
  h.b_rs = rm ? 1 : np;
  h.b_cs = rm ? np : 1;
  h.c_rs = np;
  h.c_cs = 1;
  const GemmDesc *d = cached_descs(std::vector<GemmDesc>{h});
  if (!d) return VES3D_MEMORY_ERROR;
  return gemm_grouped<double>(d, 1, 1, h.M, h.N, X, R, out, 1.0, 0.0, st,
                              gemm_pick_tile(&h, 1));
}

}  // namespace

// ------------------------------------------------------------------ MovePole
struct MovePoleG {
  int p = 0, np = 0, nc = 0, scheme = 0;  // 0 Direct, 1 ViaSpHarm, 2 DirectEagerEval
  DevBuf R;                                // [p+1][np][np] row-major
  DevBuf Rsh;                              // per latitude, per degree blocks
  std::vector<long> rsh_off;               // offset of latitude i's blocks
  std::vector<int> h_dof;                  // same-order offset of degree n (per function)
  int *dof = nullptr;
  std::vector<const double *> ops;
  std::vector<long> nsub;
  std::vector<DevBuf> shc;                 // ViaSpHarm: same-order coefficients per operand
  DevBuf work, work2, eager;
  int eager_lat = -1;
  hipStream_t st = nullptr;

  ~MovePoleG() {
    if (dof) (void)hipFree(dof);
  }

  int init(int p_, int scheme_, const double *rot_host, const double *sh_host) {
    if (p_ < 1 || scheme_ < 0 || scheme_ > 2) return VES3D_INVALID_PARAMETER;
    p = p_;
    scheme = scheme_;
    np = sh_ngrid(p);
    nc = sh_ncoef(p);
    std::vector<double> Rh;
    if (rot_host) {  // the reference's file order: column-major blocks
      Rh.resize((size_t)(p + 1) * np * np);
      for (int i = 0; i <= p; ++i)
        for (int r = 0; r < np; ++r)
          for (int c = 0; c < np; ++c)
            Rh[((size_t)i * np + r) * np + c] = rot_host[((size_t)i * np + c) * np + r];
    } else {
      build_direct_mats(p, Rh);
    }
    V3D_CHECK(R.reserve(Rh.size()));
    V3D_HIP(hipMemcpy(R.p, Rh.data(), sizeof(double) * Rh.size(), hipMemcpyHostToDevice));
    std::vector<double> S;
    if (sh_host) {  // column-major blocks in the file
      for (int i = 0; i <= p; ++i)
        for (int n = 0; n <= p; ++n) {
          const int L = deg_len(p, n);
          const size_t base = S.size();
          S.resize(base + (size_t)L * L);
          for (int r = 0; r < L; ++r)
            for (int c = 0; c < L; ++c) S[base + (size_t)r * L + c] = sh_host[base + (size_t)c * L + r];
        }
    } else {
      build_spharm_mats(p, Rh, S);
    }
    V3D_CHECK(Rsh.reserve(S.size()));
    V3D_HIP(hipMemcpy(Rsh.p, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice));
    long per_lat = 0;
    for (int n = 0; n <= p; ++n) per_lat += (long)deg_len(p, n) * deg_len(p, n);
    rsh_off.resize(p + 1);
    for (int i = 0; i <= p; ++i) rsh_off[i] = i * per_lat;
    h_dof.assign(p + 2, 0);
    for (int n = 0; n <= p; ++n) h_dof[n + 1] = h_dof[n] + deg_len(p, n);
    V3D_HIP(hipMalloc(&dof, sizeof(int) * h_dof.size()));
    V3D_HIP(hipMemcpy(dof, h_dof.data(), sizeof(int) * h_dof.size(), hipMemcpyHostToDevice));
    return VES3D_SUCCESS;
  }

  int set_operands(const double *const *arr, const size_t *ns, int num, hipStream_t s) {
    if (num < 0 || (num > 0 && (!arr || !ns))) return VES3D_INVALID_PARAMETER;
    st = s;
    ops.assign(arr, arr + num);
    nsub.assign(ns, ns + num);
    eager_lat = -1;
    if (scheme != 1) return VES3D_SUCCESS;
    shc.resize(num);
    for (int v = 0; v < num; ++v) {
      const long cnt = (long)nsub[v] * nc;
      V3D_CHECK(shc[v].reserve(cnt));
      V3D_CHECK(work.reserve(cnt));
      V3D_CHECK(ves3d_hip_shtrans_forward(p, nsub[v], ops[v], work.p, st));
      k_same_order<<<nb(cnt), 256, 0, st>>>(work.p, p, nsub[v], dof, 1, shc[v].p);
      V3D_CHECK(launch_status());
    }
    return VES3D_SUCCESS;
  }

  int direct(int i, int j, double *const *res) {
    const double *Ri = R.p + (size_t)i * np * np;
    for (size_t v = 0; v < ops.size(); ++v) {
      const long cnt = (long)nsub[v] * np;
      V3D_CHECK(work.reserve(cnt));
      k_shift_lon<<<nb(cnt), 256, 0, st>>>(ops[v], nsub[v], p, j, 1, work.p);
      V3D_CHECK(launch_status());
      V3D_CHECK(rot_gemm(work.p, nsub[v], np, Ri, true, res[v], st));
    }
    return VES3D_SUCCESS;
  }

  // all longitudes of latitude i: eager[j][v] blocks, one GEMM per operand
  int update_eager(int i) {
    const double *Ri = R.p + (size_t)i * np * np;
    const int K = 2 * p;
    size_t tot = 0;
    for (long n : nsub) tot += (size_t)K * n * np;
    V3D_CHECK(eager.reserve(tot));
    size_t off = 0;
    for (size_t v = 0; v < ops.size(); ++v) {
      const long cnt = (long)K * nsub[v] * np;
      V3D_CHECK(work.reserve(cnt));
      k_shift_lon<<<nb(cnt), 256, 0, st>>>(ops[v], nsub[v], p, 0, K, work.p);
      V3D_CHECK(launch_status());
      V3D_CHECK(rot_gemm(work.p, (long)K * nsub[v], np, Ri, true, eager.p + off, st));
      off += cnt;
    }
    eager_lat = i;
    return VES3D_SUCCESS;
  }

  int eager_copy(int j, double *const *res) {
    size_t off = 0;
    for (size_t v = 0; v < ops.size(); ++v) {
      const size_t blk = (size_t)nsub[v] * np;
      V3D_HIP(hipMemcpyAsync(res[v], eager.p + off + (size_t)j * blk, sizeof(double) * blk,
                             hipMemcpyDeviceToDevice, st));
      off += (size_t)2 * p * blk;
    }
    return VES3D_SUCCESS;
  }

  int via_spharm(int i, int j, double *const *res) {
    const double lambda = kPi / p * j;
    for (size_t v = 0; v < ops.size(); ++v) {
      const long ns = (long)nsub[v], cnt = ns * nc;
      V3D_CHECK(work.reserve(cnt));
      V3D_CHECK(work2.reserve(cnt));
      k_lon_rotate<<<nb(cnt), 256, 0, st>>>(shc[v].p, p, ns, dof, lambda, work.p);
      V3D_CHECK(launch_status());
      // per degree: work2[n] (len x ns, column-major) = M_n (len x len) work[n]
      std::vector<GemmDesc> h(p + 1);
      long moff = rsh_off[i];
      for (int n = 0; n <= p; ++n) {
        const int L = deg_len(p, n);
        GemmDesc &d = h[n];
        d = GemmDesc{};
        d.M = L;
        d.N = (int)ns;
        d.K = L;
        d.a_off = moff;
        d.a_rs = L;
        d.a_cs = 1;
        d.b_off = ns * h_dof[n];
        d.b_rs = 1;
        d.b_cs = L;
        d.c_off = ns * h_dof[n];
        d.c_rs = 1;
        d.c_cs = L;
        moff += (long)L * L;
      }
      const GemmDesc *d = cached_descs(h);
      if (!d) return VES3D_MEMORY_ERROR;
      V3D_CHECK(gemm_grouped<double>(d, p + 1, 1, 2 * p, (int)ns, Rsh.p, work.p, work2.p, 1.0,
                                     0.0, st, gemm_pick_tile(h.data(), p + 1)));
      k_same_order<<<nb(cnt), 256, 0, st>>>(work2.p, p, ns, dof, 0, work.p);
      V3D_CHECK(launch_status());
      V3D_CHECK(ves3d_hip_shtrans_backward(p, ns, work.p, 0, 0, res[v], st));
    }
    return VES3D_SUCCESS;
  }

  int apply(int i, int j, double *const *res) {
    if (i < 0 || i > p || j < 0 || j >= 2 * p || (!res && !ops.empty()))
      return VES3D_INVALID_PARAMETER;
    if (scheme == 0) return direct(i, j, res);
    if (scheme == 2) {
      if (i != eager_lat) V3D_CHECK(update_eager(i));
      return eager_copy(j, res);
    }
    return via_spharm(i, j, res);
  }
};

}  // namespace v3d

using v3d::MovePoleG;

extern "C" int ves3d_hip_movepole_create(int p, int scheme, const double *all_rot_mats_host,
                                         const double *sh_rot_mats_host, void **handle_host) {
  if (!handle_host) return VES3D_INVALID_PARAMETER;
  *handle_host = nullptr;
  MovePoleG *g = new (std::nothrow) MovePoleG();
  if (!g) return VES3D_MEMORY_ERROR;
  const int s = g->init(p, scheme, all_rot_mats_host, sh_rot_mats_host);
  if (s != VES3D_SUCCESS) {
    delete g;
    return s;
  }
  *handle_host = g;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_movepole_destroy(void *handle) {
  delete static_cast<MovePoleG *>(handle);
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_movepole_set_operands(void *handle, const double *const *arr_host,
                                               const size_t *nsub_host, int num,
                                               ves3d_stream_t stream) {
  if (!handle) return VES3D_INVALID_PARAMETER;
  return static_cast<MovePoleG *>(handle)->set_operands(arr_host, nsub_host, num,
                                                        (hipStream_t)stream);
}

extern "C" int ves3d_hip_movepole_apply(void *handle, int trg_i, int trg_j,
                                        double *const *results_host) {
  if (!handle) return VES3D_INVALID_PARAMETER;
  return static_cast<MovePoleG *>(handle)->apply(trg_i, trg_j, results_host);
}

extern "C" int ves3d_hip_movepole_matrices(void *handle, double *all_rot_mats_host) {
  if (!handle || !all_rot_mats_host) return VES3D_INVALID_PARAMETER;
  const MovePoleG *g = static_cast<MovePoleG *>(handle);
  return v3d::hip_status(hipMemcpy(all_rot_mats_host, g->R.p,
                                   sizeof(double) * (size_t)(g->p + 1) * g->np * g->np,
                                   hipMemcpyDeviceToHost));
}

// Device<>::AggregateRotation (src/DeviceCPU.cc:644-681): for every
// longitude jj and operand ii, res[n_vec jj + ii] = CircShift(mat, jj np) vec[ii]
// with mat the column-major np x np matrix of one latitude.  One GEMM per
// operand over all nlong longitudes.
extern "C" int ves3d_hip_aggregate_rotation_f64(int sh_order, int nlat, int nlong, int n_vec,
                                                const int *n_sub_host, const double *mat,
                                                const double *const *vec_host,
                                                double *const *res_host, ves3d_stream_t stream) {
  using namespace v3d;
  if (sh_order < 1 || nlat != sh_order + 1 || nlong != 2 * sh_order || n_vec < 0)
    return VES3D_INVALID_PARAMETER;
  if (n_vec == 0) return VES3D_SUCCESS;
  if (!n_sub_host || !mat || !vec_host || !res_host) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const int np = nlat * nlong;
  DevBuf work, out;
  int rc = VES3D_SUCCESS;
  for (int v = 0; v < n_vec && rc == VES3D_SUCCESS; ++v) {
    const long ns = n_sub_host[v], cnt = (long)nlong * ns * np;
    if (ns <= 0) continue;
    if ((rc = work.reserve(cnt)) != VES3D_SUCCESS) break;
    if ((rc = out.reserve(cnt)) != VES3D_SUCCESS) break;
    k_shift_lon<<<nb(cnt), 256, 0, st>>>(vec_host[v], ns, sh_order, 0, nlong, work.p);
    if ((rc = launch_status()) != VES3D_SUCCESS) break;
    if ((rc = rot_gemm(work.p, (long)nlong * ns, np, mat, false, out.p, st)) != VES3D_SUCCESS)
      break;
    for (int jj = 0; jj < nlong && rc == VES3D_SUCCESS; ++jj)
      rc = hip_status(hipMemcpyAsync(res_host[(size_t)n_vec * jj + v], out.p + (size_t)jj * ns * np,
                                     sizeof(double) * ns * np, hipMemcpyDeviceToDevice, st));
    if (rc == VES3D_SUCCESS) rc = hip_status(hipStreamSynchronize(st));
  }
  (void)hipStreamSynchronize(st);  // work / out are released on return
  return rc;
}

// shtrans.hip -- the device SHT of the reference's SHTrans<Container, Mats>
// (src/SHTrans.cc) on the order-major coefficient layout:
//   coefficient block j (order (j+1)/2; cosine for odd j and j = 0, sine for
//   even j > 0) of every function in turn, degrees ascending inside a block
//   (src/SHTrans.cc:58-120).  nfun functions, p(p+2) coefficients each.
// The transforms reuse the SphericalHarmonics GEMM plans (identical
// normalisation: the forward DFT carries 1/(2p), 1/p, the Legendre matrices
// w sqrt(2 pi) Pbar; the product equals MatFourierInv x MatLegendreInv) and
// convert between the function-major and order-major layouts.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "runtime.hpp"
#include "sh_tables.hpp"
#include "sht.hpp"

namespace v3d {

namespace {
std::mutex g_mu;
std::map<std::pair<int, int>, int2 *> g_cinfo;  // (device, p) -> per-coefficient (block offset, length)

const int2 *cinfo_for(int p) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, p);
  auto it = g_cinfo.find(key);
  if (it != g_cinfo.end()) return it->second;
  const int nc = sh_ncoef(p);
  std::vector<int2> ci(nc);
  for (int j = 0; j < 2 * p; ++j)
    for (int k = 0; k < sh_blocklen(p, j); ++k) ci[sh_blockoff(p, j) + k] = make_int2(sh_blockoff(p, j), sh_blocklen(p, j));
  int2 *d = nullptr;
  if (hipMalloc(&d, sizeof(int2) * nc) != hipSuccess) return nullptr;
  if (hipMemcpy(d, ci.data(), sizeof(int2) * nc, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  g_cinfo[key] = d;
  return d;
}

// dir = 0: function-major -> order-major; 1: the reverse.  `scale` (optional,
// per coefficient of one function) multiplies on the way: ScaleFreq.
__global__ void k_relayout(const double *__restrict__ in, long nf, int nc,
                           const int2 *__restrict__ cinfo, const double *__restrict__ scale,
                           int dir, double *__restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf * nc;
       i += (long)gridDim.x * blockDim.x) {
    const long f = i / nc;
    const int c = (int)(i % nc);
    const int2 bl = cinfo[c];
    const long o = nf * bl.x + f * bl.y + (c - bl.x);
    const double s = scale ? scale[c] : 1.0;
    if (dir == 0) out[o] = s * in[i];
    else out[i] = s * in[o];
  }
}

inline unsigned gb(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

int check(int p, size_t nfun) {
  if (p < 1 || p > 128) return VES3D_INVALID_PARAMETER;
  (void)nfun;
  return VES3D_SUCCESS;
}
}  // namespace

}  // namespace v3d

using namespace v3d;

extern "C" {

// SHTrans::forward (src/SHTrans.cc:102-120): grid [nfun][(p+1)2p] -> coefficients (order-major).
int ves3d_hip_shtrans_forward(int p, size_t nfun, const double *grid, double *shc,
                              ves3d_stream_t stream) {
  V3D_CHECK(check(p, nfun));
  if (nfun == 0) return VES3D_SUCCESS;
  if (!grid || !shc) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const long nc = sh_ncoef(p), nf = (long)nfun;
  const int2 *ci = cinfo_for(p);
  double *tmp = (double *)workspace(kWsSht1, sizeof(double) * nf * nc);
  double *w = (double *)workspace(kWsSht0, sizeof(double) * grid2shc_work(nf, p, p));
  if (!ci || !tmp || !w) return VES3D_MEMORY_ERROR;
  V3D_CHECK(sht_grid2shc<double>(grid, nf, p, p, tmp, w, st));
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(tmp, nf, (int)nc, ci, nullptr, 0, shc);
  return launch_status();
}

// SHTrans::backward / backward_du / backward_dv / backward_d2u / backward_d2v /
// backward_duv (src/SHTrans.cc:82-170): du, dv = latitude / longitude
// derivative orders (0..2, du + dv <= 2).  The latitude derivative follows the
// d1legTrans / d2legTrans matrices of the reference.
int ves3d_hip_shtrans_backward(int p, size_t nfun, const double *shc, int du, int dv,
                               double *grid, ves3d_stream_t stream) {
  V3D_CHECK(check(p, nfun));
  if (du < 0 || dv < 0 || du + dv > 2) return VES3D_INVALID_PARAMETER;
  if (nfun == 0) return VES3D_SUCCESS;
  if (!grid || !shc) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const long nc = sh_ncoef(p), nf = (long)nfun;
  const int2 *ci = cinfo_for(p);
  double *tmp = (double *)workspace(kWsSht1, sizeof(double) * nf * nc);
  double *w = (double *)workspace(kWsSht0, sizeof(double) * shc2grid_work(nf, p, p));
  if (!ci || !tmp || !w) return VES3D_MEMORY_ERROR;
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(shc, nf, (int)nc, ci, nullptr, 1, tmp);
  V3D_CHECK(launch_status());
  return sht_synth<double>(tmp, nf, p, p, du, dv, grid, w, st, du == 1 ? -1.0 : 1.0);
}

// SHTrans::ScaleFreq (src/SHTrans.cc:192-223): out = scale (per (m, n)) * in,
// scale given per coefficient of one function in the function-major order.
int ves3d_hip_shtrans_scale_freq(int p, size_t nfun, const double *shc_in, const double *scale,
                                 double *shc_out, ves3d_stream_t stream) {
  V3D_CHECK(check(p, nfun));
  if (nfun == 0) return VES3D_SUCCESS;
  if (!shc_in || !scale || !shc_out) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const long nc = sh_ncoef(p), nf = (long)nfun;
  const int2 *ci = cinfo_for(p);
  double *tmp = (double *)workspace(kWsSht1, sizeof(double) * nf * nc);
  if (!ci || !tmp) return VES3D_MEMORY_ERROR;
  // order-major -> function-major with scaling, then back
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(shc_in, nf, (int)nc, ci, scale, 1, tmp);
  V3D_CHECK(launch_status());
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(tmp, nf, (int)nc, ci, nullptr, 0, shc_out);
  return launch_status();
}

// SHTrans::lowPassFilter (src/SHTrans.cc:172-180): keeps degrees n <= freq.
int ves3d_hip_shtrans_low_pass(int p, size_t nfun, int freq, const double *in, double *out,
                               ves3d_stream_t stream) {
  V3D_CHECK(check(p, nfun));
  if (nfun == 0) return VES3D_SUCCESS;
  if (!in || !out) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const long nc = sh_ncoef(p), nf = (long)nfun;
  std::vector<double> box(nc);
  for (int j = 0; j < 2 * p; ++j)
    for (int k = 0; k < sh_blocklen(p, j); ++k)
      box[sh_blockoff(p, j) + k] = ((j + 1) / 2 + k) > freq ? 0.0 : 1.0;
  double *dbox = (double *)workspace(kWsSht2, sizeof(double) * nc);
  double *tmp = (double *)workspace(kWsSht1, sizeof(double) * 2 * nf * nc);
  double *w = (double *)workspace(kWsSht0, sizeof(double) * std::max(shc2grid_work(nf, p, p),
                                                                       grid2shc_work(nf, p, p)));
  const int2 *ci = cinfo_for(p);
  if (!dbox || !tmp || !w || !ci) return VES3D_MEMORY_ERROR;
  V3D_HIP(hipMemcpyAsync(dbox, box.data(), sizeof(double) * nc, hipMemcpyHostToDevice, st));
  V3D_CHECK(sht_grid2shc<double>(in, nf, p, p, tmp, w, st));
  // scale in the function-major layout (to order-major and back)
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(tmp, nf, (int)nc, ci, dbox, 0, tmp + nf * nc);
  V3D_CHECK(launch_status());
  k_relayout<<<gb(nf * nc), 256, 0, st>>>(tmp + nf * nc, nf, (int)nc, ci, nullptr, 1, tmp);
  V3D_CHECK(launch_status());
  V3D_CHECK(sht_shc2grid<double>(tmp, nf, p, p, out, nullptr, nullptr, w, st));
  return hip_status(hipStreamSynchronize(st));  // box is host memory
}

// Resample (src/HelperFuns.cc:436-493): grid at p -> grid at q by SH
// truncation / zero padding.
int ves3d_hip_resample(int p, int q, size_t nfun, const double *in, double *out,
                       ves3d_stream_t stream) {
  if (p < 1 || q < 1 || p > 128 || q > 255) return VES3D_INVALID_PARAMETER;
  if (nfun == 0) return VES3D_SUCCESS;
  if (!in || !out) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  const long nf = (long)nfun;
  if (p == q)
    return hip_status(hipMemcpyAsync(out, in, sizeof(double) * nf * sh_ngrid(p), hipMemcpyDeviceToDevice, st));
  const int m = std::min(p, q);
  double *tmp = (double *)workspace(kWsSht1, sizeof(double) * nf * sh_ncoef(m));
  double *w = (double *)workspace(kWsSht0, sizeof(double) * std::max(grid2shc_work(nf, p, m),
                                                                       shc2grid_work(nf, m, q)));
  if (!tmp || !w) return VES3D_MEMORY_ERROR;
  V3D_CHECK(sht_grid2shc<double>(in, nf, p, m, tmp, w, st));
  return sht_shc2grid<double>(tmp, nf, m, q, out, nullptr, nullptr, w, st);
}

}  // extern "C"

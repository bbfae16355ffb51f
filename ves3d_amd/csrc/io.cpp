// io.cpp -- the input formats on either side of the step: ASCII data files
// (DataIO::ReadDataStl, src/DataIO_templates.cc:72-98), the shape gallery +
// geometry specification initialisation (InitializeShapes,
// src/HelperFuns.cc:366-432, called by Simulation::setup_from_options,
// src/ves3d_simulation.cc:126-168) and the vesicle-properties file
// (src/ves3d_simulation.cc:174-196).  Host code; no device work.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ves3d_hip.h"

namespace {

// R = Rz(rz2) Ry(ry) Rz(rz1), row-major (src/HelperFuns.cc:314-346).
void rotation_zyz(double rz1, double ry, double rz2, double R[9]) {
  const double R1[9] = {std::cos(rz1), -std::sin(rz1), 0.0, std::sin(rz1), std::cos(rz1), 0.0,
                        0.0, 0.0, 1.0};
  const double R2[9] = {std::cos(ry), 0.0, std::sin(ry), 0.0, 1.0, 0.0, -std::sin(ry), 0.0,
                        std::cos(ry)};
  const double R3[9] = {std::cos(rz2), -std::sin(rz2), 0.0, std::sin(rz2), std::cos(rz2), 0.0,
                        0.0, 0.0, 1.0};
  double Ri[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Ri[i * 3 + j] = 0.0;
      for (int k = 0; k < 3; ++k) Ri[i * 3 + j] += R2[i * 3 + k] * R1[k * 3 + j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R[i * 3 + j] = 0.0;
      for (int k = 0; k < 3; ++k) R[i * 3 + j] += R3[i * 3 + k] * Ri[k * 3 + j];
    }
}

}  // namespace

extern "C" int ves3d_read_ascii(const char *path, double *out_host, size_t capacity,
                                size_t *count_host) {
  if (!path || !count_host) return VES3D_INVALID_PARAMETER;
  std::ifstream f(path);
  if (!f.good()) return VES3D_INVALID_PARAMETER;
  size_t n = 0;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    std::istringstream is(line);
    while (is.good()) {
      while (is.peek() == ' ' || is.peek() == '\t') is.get();
      if (is.peek() == '#' || is.peek() == EOF) break;
      double d;
      if (!(is >> d)) break;
      if (out_host) {
        if (n >= capacity) return VES3D_MEMORY_ERROR;
        out_host[n] = d;
      }
      ++n;
    }
  }
  *count_host = n;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_initialize_shapes(int p, size_t nves, const double *gallery,
                                       size_t gallery_len, const double *geo_spec,
                                       double *X_host) {
  if (p < 1 || (nves > 0 && (!gallery || !geo_spec || !X_host))) return VES3D_INVALID_PARAMETER;
  const size_t stride = 2 * (size_t)p * (p + 1);
  const size_t nshapes = gallery_len / (3 * stride);
  for (size_t v = 0; v < nves; ++v) {
    const double *g = geo_spec + 8 * v;
    const double si = g[0];
    if (si < 0 || si != std::floor(si) || (size_t)si >= nshapes) return VES3D_INVALID_PARAMETER;
    const double scale = g[1];
    const double cen[3] = {g[2], g[3], g[4]};
    double R[9];
    rotation_zyz(g[5], g[6], g[7], R);
    const double *shp = gallery + (size_t)si * 3 * stride;
    double *out = X_host + v * 3 * stride;
    for (size_t i = 0; i < stride; ++i) {
      const double xyz[3] = {shp[i] * scale, shp[stride + i] * scale, shp[2 * stride + i] * scale};
      for (int c = 0; c < 3; ++c) {
        double a = 0.0;
        for (int k = 0; k < 3; ++k) a += R[c * 3 + k] * xyz[k];
        out[c * stride + i] = a + cen[c];
      }
    }
  }
  return VES3D_SUCCESS;
}

extern "C" int ves3d_read_vesicle_props(const char *path, size_t nves, double *props_host) {
  if (!path || (nves > 0 && !props_host)) return VES3D_INVALID_PARAMETER;
  const size_t nprops = 3;  // VesicleProperties::n_props (src/VesicleProps.cc:16)
  size_t n = 0;
  int e = ves3d_read_ascii(path, nullptr, 0, &n);
  if (e != VES3D_SUCCESS) return e;
  if (n < nprops * nves) return VES3D_IO_BAD_STREAM;
  std::vector<double> all(n);
  e = ves3d_read_ascii(path, all.data(), n, &n);
  if (e != VES3D_SUCCESS) return e;
  // the file holds one row of (bending modulus, viscosity contrast, excess
  // density) per vesicle; the properties come out by column
  for (size_t v = 0; v < nves; ++v)
    for (size_t k = 0; k < nprops; ++k) props_host[k * nves + v] = all[v * nprops + k];
  return VES3D_SUCCESS;
}

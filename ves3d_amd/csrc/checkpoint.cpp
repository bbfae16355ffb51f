// checkpoint.cpp -- the reference's ASCII checkpoint: what Monitor writes
// every checkpoint stride (src/Monitor.cc:58-78) and Simulation reads back to
// restart (src/ves3d_simulation.cc:119-123, :237-249): Parameters::pack
// followed by EvolveSurface::pack, which nests VesicleProperties::pack and
// Surface::pack -> Vectors::pack -> Scalars::pack -> Array::pack.  Every
// block is "KEY\n", "key: value" lines and "/KEY\n"; numbers are written with
// std::scientific and 16 digits (Monitor sets both on the stream), which
// round-trips every double.  The reader accepts the layouts the reference's
// unpack methods accept: version <= 590 (no name lines in Scalars/Vectors,
// vesicle properties not stored), < 592 (six deprecated solver keys), < 593
// (init_file_name / cntrs_file_name instead of the gallery and geometry
// file keys).  Host code; no device work.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ves3d_hip.h"

namespace {

// ---- writing -------------------------------------------------------------
struct Out {
  std::ostringstream os;
  void key(const char *k) { os << k << "\n"; }
  void kv(const char *k, long v) { os << k << ": " << v << "\n"; }
  void kvd(const char *k, double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.16e", v);
    os << k << ": " << b << "\n";
  }
  void kvs(const char *k, const char *v) { os << k << ": " << v << "\n"; }
  // string parameter lines end in " |" so that an empty value does not make
  // the reader consume the next line (src/Parameters.cc:458-462)
  void kvbar(const char *k, const char *v) { os << k << ": " << v << " |\n"; }
  void array(const char *name, const double *v, size_t n, const char *version) {
    key("ARRAY");
    kvs("version", version);
    kvs("name", name);
    os << "size: " << n << "\n";
    os << "data: ";
    char b[64];
    for (size_t i = 0; i < n; ++i) {
      std::snprintf(b, sizeof(b), "%.16e", v[i]);
      os << b << (i + 1 < n ? " " : "");
    }
    os << "\n/ARRAY\n";
  }
};

// ---- reading -------------------------------------------------------------
struct In {
  std::istringstream is;
  bool ok = true;
  std::string tok() {
    std::string s;
    if (!(is >> s)) ok = false;
    return s;
  }
  bool expect(const char *s) {
    const std::string t = tok();
    if (t != s) ok = false;
    return ok;
  }
  long geti() {
    const std::string t = tok();
    char *e = nullptr;
    const long v = std::strtol(t.c_str(), &e, 10);
    if (e == t.c_str()) ok = false;
    return v;
  }
  double getd() {
    const std::string t = tok();
    char *e = nullptr;
    const double v = std::strtod(t.c_str(), &e);
    if (e == t.c_str()) ok = false;
    return v;
  }
  // "version: N [+]" (the optional key is consumed by the caller's next
  // read); returns -1 when the block has no version line
  int version(std::string &next) {
    next = tok();
    if (next != "version:") return -1;
    int v = (int)geti();
    next = tok();
    if (next == "+") {
      ++v;
      next = tok();
    }
    return v;
  }
  // a string parameter written by kvbar: the value, or "" when it is "|"
  std::string barred() {
    std::string s = tok();
    if (s == "|") return std::string();
    expect("|");
    return s;
  }
};

void copy_str(char *dst, size_t cap, const std::string &s) {
  std::snprintf(dst, cap, "%s", s.c_str());
}

// one ARRAY block into out (resized to its size)
bool read_array(In &in, std::vector<double> &out) {
  if (!in.expect("ARRAY")) return false;
  std::string key;
  in.version(key);
  if (key != "name:") return false;
  in.tok();  // the array's name
  if (!in.expect("size:")) return false;
  const long n = in.geti();
  if (n < 0 || !in.expect("data:")) return false;
  out.resize((size_t)n);
  for (long i = 0; i < n && in.ok; ++i) out[i] = in.getd();
  return in.ok && in.expect("/ARRAY");
}

}  // namespace

extern "C" int ves3d_checkpoint_defaults(int sh_order, size_t nves,
                                         ves3d_checkpoint_params *prm) {
  if (!prm || sh_order < 1) return VES3D_INVALID_PARAMETER;
  std::memset(prm, 0, sizeof(*prm));
  // Parameters<T>::init defaults (src/Parameters.cc:23-58) with the order,
  // count and the implicit-step choices of this library
  prm->version = 600;
  prm->n_surfs = (int)nves;
  prm->sh_order = sh_order;
  prm->filter_freq = 2 * sh_order / 3;
  prm->upsample_freq = 2 * sh_order;
  prm->bending_modulus = 1e-2;
  prm->viscosity_contrast = 1.0;
  prm->time_horizon = 1.0;
  prm->ts = 1e-3;
  prm->time_tol = 1e-6;
  prm->time_iter_max = 100;
  prm->solve_for_velocity = 1;
  copy_str(prm->scheme, sizeof(prm->scheme), "GloballyImplicit");
  copy_str(prm->time_precond, sizeof(prm->time_precond), "DiagonalSpectral");
  copy_str(prm->bg_flow, sizeof(prm->bg_flow), "ShearFlow");
  copy_str(prm->singular_stokes, sizeof(prm->singular_stokes), "ViaSpHarm");
  copy_str(prm->rep_type, sizeof(prm->rep_type), "PolyKReparam");
  prm->rep_maxit = 10;
  prm->rep_filter_freq = sh_order;
  prm->rep_ts = -1.0;
  prm->rep_tol = 1e-4;
  prm->rep_exponent = 4.0;
  prm->repul_dist = 5e-2;
  prm->bg_flow_param = 1e-1;
  prm->periodic_length = -1.0;
  prm->checkpoint_stride = -1.0;
  prm->error_factor = 1.0;
  prm->num_threads = -1;
  prm->gravity_field[2] = -1.0;
  prm->sht_resample_order = 2 * sh_order;
  prm->diff_filter_freq = 2 * sh_order / 3;
  prm->reparam_filter_freq = sh_order;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_checkpoint_write(const char *path, const ves3d_checkpoint_params *prm,
                                      size_t nves, const double *X, const double *bending_modulus,
                                      const double *viscosity_contrast,
                                      const double *excess_density) {
  if (!path || !prm || !X || prm->sh_order < 1) return VES3D_INVALID_PARAMETER;
  if (nves > 0 && (!bending_modulus || !viscosity_contrast || !excess_density))
    return VES3D_INVALID_PARAMETER;
  const int p = prm->sh_order;
  const long stride = 2L * p * (p + 1);
  char ver[32];
  std::snprintf(ver, sizeof(ver), "%d", prm->version);
  Out o;
  // Parameters<T>::pack (src/Parameters.cc:424-470)
  o.key("PARAMETERS");
  o.kvs("version", ver);
  o.kvs("name", "parameters");
  o.kv("n_surfs", prm->n_surfs);
  o.kv("sh_order", prm->sh_order);
  o.kv("filter_freq", prm->filter_freq);
  o.kv("upsample_freq", prm->upsample_freq);
  o.kvd("bending_modulus", prm->bending_modulus);
  o.kvd("viscosity_contrast", prm->viscosity_contrast);
  o.kvd("time_horizon", prm->time_horizon);
  o.kvd("ts", prm->ts);
  o.kvd("time_tol", prm->time_tol);
  o.kv("time_iter_max", prm->time_iter_max);
  o.kv("time_adaptive", prm->time_adaptive ? 1 : 0);
  o.kv("solve_for_velocity", prm->solve_for_velocity ? 1 : 0);
  o.kv("pseudospectral", prm->pseudospectral ? 1 : 0);
  o.kvs("scheme", prm->scheme);
  o.kvs("time_precond", prm->time_precond);
  o.kvs("bg_flow", prm->bg_flow);
  o.kvs("singular_stokes", prm->singular_stokes);
  o.kvs("rep_type", prm->rep_type);
  o.kv("rep_maxit", prm->rep_maxit);
  o.kv("rep_upsample", prm->rep_upsample ? 1 : 0);
  o.kv("rep_filter_freq", prm->rep_filter_freq);
  o.kvd("rep_ts", prm->rep_ts);
  o.kvd("rep_tol", prm->rep_tol);
  o.kvd("rep_exponent", prm->rep_exponent);
  o.kvd("repul_dist", prm->repul_dist);
  o.kvd("bg_flow_param", prm->bg_flow_param);
  o.kvd("periodic_length", prm->periodic_length);
  o.kv("interaction_upsample", prm->interaction_upsample ? 1 : 0);
  o.kv("checkpoint", prm->checkpoint ? 1 : 0);
  o.kvd("checkpoint_stride", prm->checkpoint_stride);
  o.kvbar("write_vtk", "");
  o.kvbar("shape_gallery_file", "");
  o.kvbar("vesicle_geometry_file", "");
  o.kvbar("vesicle_props_file", "");
  o.kvbar("checkpoint_file_name", "");
  o.kvbar("load_checkpoint", "");
  o.kvd("error_factor", prm->error_factor);
  o.kv("num_threads", prm->num_threads);
  o.kvd("excess_density", prm->excess_density);
  {
    char b[200];
    std::snprintf(b, sizeof(b), "gravity_field: %.16e %.16e %.16e\n", prm->gravity_field[0],
                  prm->gravity_field[1], prm->gravity_field[2]);
    o.os << b;
  }
  o.key("/PARAMETERS");
  // EvolveSurface::pack (src/EvolveSurface.cc:477-490)
  o.key("EVOLVE");
  o.kvs("version", ver);
  o.kvs("name", "evolve_surface");
  // VesicleProperties::pack (src/VesicleProps.cc:110-122)
  o.key("VESICLEPROPS");
  o.kvs("version", ver);
  o.kvs("name", "vesicle_properties");
  o.array("bending_modulus", bending_modulus, nves, ver);
  o.array("viscosity_contrast", viscosity_contrast, nves, ver);
  o.array("excess_density", excess_density, nves, ver);
  o.key("/VESICLEPROPS");
  // Surface::pack (src/Surface.cc:609-627)
  o.key("SURFACE");
  o.kvs("version", ver);
  o.kvs("name", "surface");
  o.kv("sh_order", p);
  o.kv("SHT_order", p);
  o.kv("SHT_rep_order", p);
  o.kvd("SHT_rep_exponent", prm->rep_exponent);
  o.kv("SHT_resample_order", prm->sht_resample_order);
  o.kv("diff_filter_freq", prm->diff_filter_freq);
  o.kv("reparam_filter_freq", prm->reparam_filter_freq);
  o.kvs("reparam_type", prm->rep_type);
  // Vectors::pack (src/Vectors.cc:86-98), Scalars::pack (src/Scalars.cc:220-236)
  o.key("VECTORS");
  o.kvs("version", ver);
  o.kvs("CoordinateOrder", "AxisMajor");
  o.key("SCALARS");
  o.kvs("version", ver);
  o.kv("nsubs", (long)nves);
  o.kv("shorder", p);
  o.kv("gridx", p + 1);
  o.kv("gridy", 2 * p);
  o.kv("stride", stride);
  o.array("position", X, nves * 3 * (size_t)stride, ver);
  o.key("/SCALARS");
  o.key("/VECTORS");
  o.key("/SURFACE");
  o.key("/EVOLVE");
  std::ofstream f(path);
  if (!f.good()) return VES3D_IO_ERROR;
  f << o.os.str();
  return f.good() ? VES3D_SUCCESS : VES3D_IO_ERROR;
}

extern "C" int ves3d_checkpoint_read(const char *path, ves3d_checkpoint_params *prm,
                                     size_t *nves_host, double *X, size_t x_capacity,
                                     double *bending_modulus, double *viscosity_contrast,
                                     double *excess_density, size_t props_capacity) {
  if (!path || !prm || !nves_host) return VES3D_INVALID_PARAMETER;
  std::ifstream f(path);
  if (!f.good()) return VES3D_IO_ERROR;
  std::stringstream buf;
  buf << f.rdbuf();
  In in;
  in.is.str(buf.str());
  std::memset(prm, 0, sizeof(*prm));
  std::string key;
  // Parameters<T>::unpack (src/Parameters.cc:476-586)
  if (!in.expect("PARAMETERS")) return VES3D_IO_BAD_STREAM;
  int version = in.version(key);
  prm->version = version;
  if (version > 590) {
    if (key != "name:") return VES3D_IO_BAD_STREAM;
    in.tok();
    key = in.tok();
  }
  if (key != "n_surfs:") return VES3D_IO_BAD_STREAM;
  prm->n_surfs = (int)in.geti();
  auto want = [&](const char *k) { return in.expect(k); };
  if (!want("sh_order:")) return VES3D_IO_BAD_STREAM;
  prm->sh_order = (int)in.geti();
  if (!want("filter_freq:")) return VES3D_IO_BAD_STREAM;
  prm->filter_freq = (int)in.geti();
  if (!want("upsample_freq:")) return VES3D_IO_BAD_STREAM;
  prm->upsample_freq = (int)in.geti();
  if (!want("bending_modulus:")) return VES3D_IO_BAD_STREAM;
  prm->bending_modulus = in.getd();
  if (!want("viscosity_contrast:")) return VES3D_IO_BAD_STREAM;
  prm->viscosity_contrast = in.getd();
  if (version < 592) {
    for (const char *k : {"position_solver_iter:", "tension_solver_iter:",
                          "position_solver_restart:", "tension_solver_restart:",
                          "position_solver_tol:", "tension_solver_tol:"}) {
      if (!want(k)) return VES3D_IO_BAD_STREAM;
      in.tok();
    }
  }
  if (!want("time_horizon:")) return VES3D_IO_BAD_STREAM;
  prm->time_horizon = in.getd();
  if (!want("ts:")) return VES3D_IO_BAD_STREAM;
  prm->ts = in.getd();
  if (!want("time_tol:")) return VES3D_IO_BAD_STREAM;
  prm->time_tol = in.getd();
  if (!want("time_iter_max:")) return VES3D_IO_BAD_STREAM;
  prm->time_iter_max = (int)in.geti();
  if (!want("time_adaptive:")) return VES3D_IO_BAD_STREAM;
  prm->time_adaptive = (int)in.geti();
  if (!want("solve_for_velocity:")) return VES3D_IO_BAD_STREAM;
  prm->solve_for_velocity = (int)in.geti();
  if (!want("pseudospectral:")) return VES3D_IO_BAD_STREAM;
  prm->pseudospectral = (int)in.geti();
  if (!want("scheme:")) return VES3D_IO_BAD_STREAM;
  copy_str(prm->scheme, sizeof(prm->scheme), in.tok());
  if (!want("time_precond:")) return VES3D_IO_BAD_STREAM;
  copy_str(prm->time_precond, sizeof(prm->time_precond), in.tok());
  if (!want("bg_flow:")) return VES3D_IO_BAD_STREAM;
  copy_str(prm->bg_flow, sizeof(prm->bg_flow), in.tok());
  if (!want("singular_stokes:")) return VES3D_IO_BAD_STREAM;
  copy_str(prm->singular_stokes, sizeof(prm->singular_stokes), in.tok());
  if (!want("rep_type:")) return VES3D_IO_BAD_STREAM;
  copy_str(prm->rep_type, sizeof(prm->rep_type), in.tok());
  if (!want("rep_maxit:")) return VES3D_IO_BAD_STREAM;
  prm->rep_maxit = (int)in.geti();
  if (!want("rep_upsample:")) return VES3D_IO_BAD_STREAM;
  prm->rep_upsample = (int)in.geti();
  if (!want("rep_filter_freq:")) return VES3D_IO_BAD_STREAM;
  prm->rep_filter_freq = (int)in.geti();
  if (!want("rep_ts:")) return VES3D_IO_BAD_STREAM;
  prm->rep_ts = in.getd();
  if (!want("rep_tol:")) return VES3D_IO_BAD_STREAM;
  prm->rep_tol = in.getd();
  if (!want("rep_exponent:")) return VES3D_IO_BAD_STREAM;
  prm->rep_exponent = in.getd();
  if (!want("repul_dist:")) return VES3D_IO_BAD_STREAM;
  prm->repul_dist = in.getd();
  if (!want("bg_flow_param:")) return VES3D_IO_BAD_STREAM;
  prm->bg_flow_param = in.getd();
  if (!want("periodic_length:")) return VES3D_IO_BAD_STREAM;
  prm->periodic_length = in.getd();
  if (!want("interaction_upsample:")) return VES3D_IO_BAD_STREAM;
  prm->interaction_upsample = (int)in.geti();
  if (!want("checkpoint:")) return VES3D_IO_BAD_STREAM;
  prm->checkpoint = (int)in.geti();
  if (!want("checkpoint_stride:")) return VES3D_IO_BAD_STREAM;
  prm->checkpoint_stride = in.getd();
  if (!want("write_vtk:")) return VES3D_IO_BAD_STREAM;
  in.barred();
  if (version < 593) {
    if (!want("init_file_name:")) return VES3D_IO_BAD_STREAM;
    in.barred();
    if (!want("cntrs_file_name:")) return VES3D_IO_BAD_STREAM;
    in.barred();
  } else {
    if (!want("shape_gallery_file:")) return VES3D_IO_BAD_STREAM;
    in.barred();
    if (!want("vesicle_geometry_file:")) return VES3D_IO_BAD_STREAM;
    in.barred();
  }
  if (version > 590) {
    if (!want("vesicle_props_file:")) return VES3D_IO_BAD_STREAM;
    in.barred();
  }
  if (!want("checkpoint_file_name:")) return VES3D_IO_BAD_STREAM;
  in.barred();
  if (!want("load_checkpoint:")) return VES3D_IO_BAD_STREAM;
  in.barred();
  if (!want("error_factor:")) return VES3D_IO_BAD_STREAM;
  prm->error_factor = in.getd();
  if (!want("num_threads:")) return VES3D_IO_BAD_STREAM;
  prm->num_threads = (int)in.geti();
  if (!want("excess_density:")) return VES3D_IO_BAD_STREAM;
  prm->excess_density = in.getd();
  if (!want("gravity_field:")) return VES3D_IO_BAD_STREAM;
  for (int k = 0; k < 3; ++k) prm->gravity_field[k] = in.getd();
  if (!want("/PARAMETERS") || !in.ok) return VES3D_IO_BAD_STREAM;

  // EvolveSurface::unpack (src/EvolveSurface.cc:494-521): the version line
  // is mandatory here
  if (!want("EVOLVE")) return VES3D_IO_BAD_STREAM;
  const int ev = in.version(key);
  if (ev < 0 || key != "name:") return VES3D_IO_BAD_STREAM;
  in.tok();
  std::vector<double> bm, vc, ed;
  if (ev > 590) {
    // VesicleProperties::unpack (src/VesicleProps.cc:124-150)
    if (!want("VESICLEPROPS")) return VES3D_IO_BAD_STREAM;
    in.version(key);
    if (key != "name:") return VES3D_IO_BAD_STREAM;
    in.tok();
    if (!read_array(in, bm) || !read_array(in, vc) || !read_array(in, ed)) return VES3D_IO_BAD_STREAM;
    if (!want("/VESICLEPROPS")) return VES3D_IO_BAD_STREAM;
  }
  // Surface::unpack (src/Surface.cc:630-686)
  if (!want("SURFACE")) return VES3D_IO_BAD_STREAM;
  in.version(key);
  if (key != "name:") return VES3D_IO_BAD_STREAM;
  in.tok();
  if (!want("sh_order:")) return VES3D_IO_BAD_STREAM;
  in.geti();
  if (!want("SHT_order:")) return VES3D_IO_BAD_STREAM;
  const long sht_order = in.geti();
  if (!want("SHT_rep_order:")) return VES3D_IO_BAD_STREAM;
  in.geti();
  if (!want("SHT_rep_exponent:")) return VES3D_IO_BAD_STREAM;
  in.getd();
  if (!want("SHT_resample_order:")) return VES3D_IO_BAD_STREAM;
  prm->sht_resample_order = (int)in.geti();
  if (!want("diff_filter_freq:")) return VES3D_IO_BAD_STREAM;
  prm->diff_filter_freq = (int)in.geti();
  if (!want("reparam_filter_freq:")) return VES3D_IO_BAD_STREAM;
  prm->reparam_filter_freq = (int)in.geti();
  if (!want("reparam_type:")) return VES3D_IO_BAD_STREAM;
  in.tok();
  // Vectors::unpack (src/Vectors.cc:100-127)
  if (!want("VECTORS")) return VES3D_IO_BAD_STREAM;
  int vv = in.version(key);
  if (vv <= 590) {  // a name line before CoordinateOrder (also with no version line)
    if (key != "name:") return VES3D_IO_BAD_STREAM;
    in.tok();
    key = in.tok();
  }
  if (key != "CoordinateOrder:") return VES3D_IO_BAD_STREAM;
  if (in.tok() != "AxisMajor") return VES3D_NOT_IMPLEMENTED;
  // Scalars::unpack (src/Scalars.cc:238-283)
  if (!want("SCALARS")) return VES3D_IO_BAD_STREAM;
  int sv = in.version(key);
  if (sv <= 590) {
    if (key != "name:") return VES3D_IO_BAD_STREAM;
    in.tok();
    key = in.tok();
  }
  if (key != "nsubs:") return VES3D_IO_BAD_STREAM;
  const long nsubs = in.geti();
  if (!want("shorder:")) return VES3D_IO_BAD_STREAM;
  const long shorder = in.geti();
  if (!want("gridx:")) return VES3D_IO_BAD_STREAM;
  const long gx = in.geti();
  if (!want("gridy:")) return VES3D_IO_BAD_STREAM;
  const long gy = in.geti();
  if (!want("stride:")) return VES3D_IO_BAD_STREAM;
  const long stride = in.geti();
  if (!in.ok || shorder != sht_order || gx != shorder + 1 || gy != 2 * shorder ||
      stride != gx * gy || nsubs < 0)
    return VES3D_IO_BAD_STREAM;
  std::vector<double> pos;
  if (!read_array(in, pos)) return VES3D_IO_BAD_STREAM;
  if (pos.size() != (size_t)nsubs * 3 * (size_t)stride) return VES3D_IO_BAD_STREAM;
  for (const char *k : {"/SCALARS", "/VECTORS", "/SURFACE", "/EVOLVE"})
    if (!want(k)) return VES3D_IO_BAD_STREAM;
  if (ev > 590 && (bm.size() != (size_t)nsubs || vc.size() != (size_t)nsubs ||
                   ed.size() != (size_t)nsubs))
    return VES3D_IO_BAD_STREAM;
  *nves_host = (size_t)nsubs;
  if (X) {
    if (x_capacity < pos.size()) return VES3D_MEMORY_ERROR;
    std::memcpy(X, pos.data(), pos.size() * sizeof(double));
  }
  // versions <= 590 store no vesicle properties: the reference takes them
  // from the parameters (VesicleProperties::setFromParams); so does this
  if (bending_modulus || viscosity_contrast || excess_density) {
    if (props_capacity < (size_t)nsubs) return VES3D_MEMORY_ERROR;
    for (long v = 0; v < nsubs; ++v) {
      if (bending_modulus) bending_modulus[v] = ev > 590 ? bm[v] : prm->bending_modulus;
      if (viscosity_contrast) viscosity_contrast[v] = ev > 590 ? vc[v] : prm->viscosity_contrast;
      if (excess_density) excess_density[v] = ev > 590 ? ed[v] : prm->excess_density;
    }
  }
  return VES3D_SUCCESS;
}

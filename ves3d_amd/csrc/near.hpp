// near.hpp -- near-singular correction of the Stokes operator
// (NearSingular<Real>, src/NearSingular.cc) on the GPU, for single- and
// double-layer densities, at the vesicles' own surface points
// (trg_is_surf) or at targets off the surfaces.
//
// Sources are the upsampled surfaces with their two poles first,
// point-major [nv][M][3] with M = 2 + 2 p_up (p_up + 1).  Surface targets
// are the p-grid surface points, point-major [nv * ng][3]; a near pair is a
// target t and a vesicle v other than t's own with a source point closer
// than r_near, paired with v's closest point k (src/NearSingular.cc:797-918).
// Targets off the surfaces belong to no vesicle and pair with every vesicle
// they come near.
#pragma once
#include <vector>

#include "common.hpp"

namespace v3d {

class NearG {
 public:
  // r_near_override <= 0 selects the reference's r_near = r_ves 2/sqrt(p_up).
  // trg_is_surf false: targets off the surfaces, searched in groups of
  // trg_per_group.
  int init(int p, int p_up, long nv, double repul_dist, double r_near_override,
           bool trg_is_surf = true, long trg_per_group = 0);
  // Finds the near pairs and projects each target on its vesicle's
  // quadratic patch; computes the per-target repulsion force (surface
  // targets).  n_trg: the number of targets off the surfaces (surface
  // targets: nv ng).  Synchronises the stream (the pair count sizes the
  // later launches).
  int setup(const double *src, const double *trg, hipStream_t st, long n_trg = 0);
  // f[nv][3][ng] += repulsion force (NearSingular::ForceRepul).
  int add_repulsion(double *f, hipStream_t st) const;
  // vel += near-singular velocity - direct sum of v at its near targets
  // (vel: [nv][3][ng] for surface targets, [n_trg][3] otherwise).
  // q: SL density times quadrature weight [nv][M][3] or null; qd: DL
  // density [nv][M][6] = (density times quadrature weight, normal) or null;
  // vup / vpole: the self velocity on the upsampled grid [nv][3][M-2] and
  // at the poles [nv][3][2]; dup / dpole: the DL density there (with qd).
  int apply(const double *q, const double *qd, const double *vup, const double *vpole,
            const double *dup, const double *dpole, double *vel, hipStream_t st);
  ~NearG();

  int p = 0, pu = 0;
  long nv = 0, ng = 0, M = 0;
  bool trg_is_surf = true;
  long tpg = 0, n_trg = 0;   // targets per search group, targets
  double repul = 0, r_near_fixed = 0, r_near = 0;
  long npairs = 0, ntrg_near = 0, ntiles = 0;

 private:
  template <typename T>
  struct Buf {
    T *p = nullptr;
    size_t n = 0;
    int reserve(size_t count);
    ~Buf();
  };
  const double *src_ = nullptr;  // the setup's sources (owned by the caller)
  Buf<double> cen, rs2, rt2, tcen, mint, mfix, ipts, wts, pxy, dvel, rep, corr;
  Buf<unsigned char> flags;
  Buf<int> cand, near_k, cnt, off, pairs, tiles, tgt, tptr, perm;
  std::vector<int> h_cand, h_cnt, h_off, h_pairs;
};

}  // namespace v3d

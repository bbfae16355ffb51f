// near.hpp -- near-singular correction of the surface Stokes operator
// (NearSingular<Real>, src/NearSingular.cc) on the GPU, for surface targets
// and a single-layer density.
//
// Sources are the upsampled surfaces with their two poles first,
// point-major [nv][M][3] with M = 2 + 2 p_up (p_up + 1); targets are the
// p-grid surface points, point-major [nv * ng][3].  A near pair is a target t
// and a vesicle v other than t's own with a source point closer than r_near,
// paired with v's closest point k (src/NearSingular.cc:797-918).
#pragma once
#include <vector>

#include "common.hpp"

namespace v3d {

class NearG {
 public:
  // r_near_override <= 0 selects the reference's r_near = r_ves 2/sqrt(p_up).
  int init(int p, int p_up, long nv, double repul_dist, double r_near_override);
  // Finds the near pairs and projects each target on its vesicle's
  // quadratic patch; computes the per-target repulsion force.  Synchronises
  // the stream (the pair count sizes the later launches).
  int setup(const double *src, const double *trg, hipStream_t st);
  // f[nv][3][ng] += repulsion force (NearSingular::ForceRepul).
  int add_repulsion(double *f, hipStream_t st) const;
  // vel[nv][3][ng] += near-singular velocity - direct sum of v at its near
  // targets.  q: SL density times quadrature weight [nv][M][3]; vup / vpole:
  // the self velocity on the upsampled grid [nv][3][M-2] and at the poles
  // [nv][3][2].
  int apply(const double *q, const double *vup, const double *vpole, double *vel,
            hipStream_t st);
  ~NearG();

  int p = 0, pu = 0;
  long nv = 0, ng = 0, M = 0;
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
  Buf<double> cen, rs2, rt2, mint, mfix, ipts, wts, pxy, dvel, rep, corr;
  Buf<unsigned char> flags;
  Buf<int> cand, near_k, cnt, off, pairs, tiles, tgt, tptr, perm;
  std::vector<int> h_cand, h_cnt, h_off, h_pairs;
};

}  // namespace v3d

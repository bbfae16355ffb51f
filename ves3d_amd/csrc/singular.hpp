// singular.hpp -- rotations, singular self-interaction build and apply.
#pragma once
#include "common.hpp"

namespace v3d {

template <typename T>
int rotate_all(const T *S, long Nv, int p, int dof, T *out, T *work, hipStream_t st);
// gathered: the input rows are already in degree-major order (the fused
// pole stage writes them so); work then needs one array of the input's size.
template <typename T>
int rotate_transpose(const T *in, long Nv, int p, int dof, T *out, T *work, hipStream_t st,
                     bool gathered = false);
// X0 [Nv][3][ngrid(p0)] FP64 -> SL/DL [Nv][3Nc][3Nc] (either may be null).
template <typename T>
int singular_build(const double *X0, long Nv, int p0, int p1, T *SL, T *DL, hipStream_t st);
// V[v] = F[v] SL[v] + G[v] DL[v].  sym: the operators' 3 x 3 blocks satisfy
// B_ab = B_ba, as singular_build writes them; only the blocks a <= b are read.
template <typename TM>
int self_apply(const TM *SL, const double *F, const TM *DL, const double *G, long Nv, int p,
               double *V, hipStream_t st, bool sym = false);

}  // namespace v3d

// gemm.hpp -- grouped, strided, batched GEMM used by every SHT / rotation stage.
//
// One launch runs a list of GEMM descriptors that share a batch count.  Each
// descriptor addresses its operands as offsets from three base pointers passed
// at launch (so the descriptor lists of a transform plan are built once and
// stay on the device) and arbitrary row/column/batch element strides (so the
// transposes and block re-layouts of the reference's SHT stages are folded into
// the operand addressing instead of being separate passes):
//     C[b](r,c) = alpha * sum_k A[b](r,k) B[b](k,c) + beta * C[b](r,c)
//     X[b](r,c) = Xbase[x_off + b*x_bs + r*x_rs + c*x_cs]
// Grid: x = output tiles (max over descriptors), y = descriptor, z = batch.
#pragma once
#include "common.hpp"

namespace v3d {

struct GemmDesc {
  int M, N, K;
  int _pad;
  long a_off, a_rs, a_cs, a_bs;
  long b_off, b_rs, b_cs, b_bs;
  long c_off, c_rs, c_cs, c_bs;
};

// Launches the descriptor list `d_descs` (device memory, `ndesc` entries) over
// `batch` batches.  `max_m`, `max_n` bound the descriptors' M and N.  `tile`
// selects the MFMA output tile (gemm_pick_tile); -1 is the 64x64 tile.
template <typename T>
int gemm_grouped(const GemmDesc *d_descs, int ndesc, int batch, int max_m,
                 int max_n, const T *A, const T *B, T *C, T alpha, T beta,
                 hipStream_t st, int tile = -1);

// MFMA output tile for a descriptor list: the shape (64 or 128 rows x 16..64
// columns) that minimises the padded multiply-add count plus operand traffic
// and a per-tile overhead over the list, with flags recording whether the
// list's A operand can be read as aligned 16-byte vectors.  The lists of the
// SH / rotation stages are skinny (N and K of 1..64, M up to 1e7), where one
// 64x64 tile pads most of its work.
int gemm_pick_tile(const GemmDesc *h, int n);

// Uploads a host descriptor list to a new device array.
GemmDesc *upload_descs(const GemmDesc *h, int n);

}  // namespace v3d

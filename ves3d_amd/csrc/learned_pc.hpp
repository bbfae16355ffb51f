// learned_pc.hpp -- a per-vesicle network in SH-coefficient space used as a
// (right) preconditioner of the implicit BiCGStab solve, trained online on
// (A e, e) pairs the solver produces.  No counterpart in the reference: it
// occupies the preconditioner slot of ImplicitPrecond
// (src/InterfacialVelocity.cc:1054-1109).
//
// For the Galerkin vector r of a step, with D the diagonal spectral
// preconditioner and, per vesicle v, z_v = (D r)_v (velocity then tension
// coefficients, 4 p(p+2) values) and s_v = |z_v|:
//     P(r)_v = z_v + s_v * g(z_v / s_v),
// g a 3-layer MLP (nin -> width -> width -> nin, tanh-GELU, FP32 weights,
// FP32 or BF16 activations).  g is trained so that P(A e) ~ e, i.e. on
// inputs z/s and targets (e_v - z_v)/s.  The solve's recursive and true
// residuals decide whether the result is used (the stepper falls back to
// the diagonal preconditioner).
#pragma once
#include <map>
#include <tuple>

#include "common.hpp"
#include "gemm.hpp"

namespace v3d {

class LearnedPC {
 public:
  ~LearnedPC();
  int init(int nin, int width, long capacity, unsigned long long seed);
  long num_params() const { return np_; }

  // ---- training interface (all device pointers, FP32 rows [B][nin])
  int forward(const float *Z, long B, float *Y, hipStream_t st);
  // loss = sum((Y-T)^2) / (2 B nin); gradients into the internal G buffer
  int loss_grad(const float *Z, const float *T, long B, double *loss_host, hipStream_t st);
  int adam(float lr, hipStream_t st);
  // Adam over shuffled minibatches with a cosine learning-rate decay on nine
  // tenths of the samples; keeps the parameters with the lowest loss on the
  // held-out tenth; *last_loss_host is that loss
  int train(int epochs, int batch, float lr, double *last_loss_host, hipStream_t st);
  // Y = (g(Z) - g(-Z)) / 2, the odd part of the network (what the solver applies)
  int odd_forward(const float *Z, long B, float *Y, hipStream_t st);
  // loss of the odd part over all recorded samples and the mean square of
  // their targets
  int dataset_loss(long have, double *loss_host, double *target_ms_host, hipStream_t st);
  // the same over the samples idx[0..m) (device indices)
  int subset_loss(const int *idx, long m, double *loss_host, double *target_ms_host,
                  hipStream_t st);
  int best_epoch() const { return best_epoch_; }
  long stored() const { return count_ < cap_ ? count_ : cap_; }

  // ---- solver interface on the Galerkin (order-major) layout
  // y holds D r on entry and P(r) on exit; the correction leaves the tension
  // coefficients of degree > tcut (the divergence filter) untouched.  frames:
  // per-vesicle 3x3 canonical frames (row-major, columns the axes) in which
  // the network sees and returns the velocity blocks; nullptr: identity.
  int apply_inplace(double *y, long nv, int p, int tcut, const int2 *cinfo, const double *frames,
                    hipStream_t st);
  // records samples from (Dr, e) in the canonical frames: inputs F^T Dr_v/s_v,
  // targets F^T (e_v - Dr_v)/s_v
  int record(const double *Dr, const double *e, long nv, int p, const int2 *cinfo,
             const double *frames, hipStream_t st);

  float *params() const { return P_; }
  float *grads() const { return G_; }
  long samples() const { return count_; }
  bool ready() const { return trained_; }
  // decoupled weight decay of the Adam steps (0: plain Adam)
  void set_weight_decay(float wd) { wd_ = wd; }
  // zeroes the Adam moments and step count (new parameters start afresh)
  int reset_optimizer();
  void set_ready(bool r) { trained_ = r; }
  // the parameters were written from outside (the BF16 weight copy is stale)
  void params_changed() { wdirty_ = true; }
  // 0: FP32 activations on FP32 MFMA tiles; 1: BF16 operands of every
  // product (activations, and the weights rounded from the FP32 master copy
  // after each update) on BF16 MFMA with FP32 accumulation; pre-activations,
  // gradients and the optimizer stay FP32.  Needs nin and width multiples
  // of 8.
  int set_activations(int mode);
  int activations() const { return act_; }

 private:
  int ensure_batch(long B);
  int refresh_bf16(hipStream_t st);
  int gemm(int M, int N, int K, const float *A, long ars, long acs, const float *Bm, long brs,
           long bcs, float *C, hipStream_t st);

  int nin_ = 0, width_ = 0;
  long np_ = 0, cap_ = 0, count_ = 0, head_ = 0, bcap_ = 0;
  long oW1_, ob1_, oW2_, ob2_, oW3_, ob3_;
  float *P_ = nullptr, *G_ = nullptr, *M_ = nullptr, *V_ = nullptr;
  float *Zs_ = nullptr, *Ts_ = nullptr;                   // sample ring buffers
  float *A1_ = nullptr, *H1_ = nullptr, *A2_ = nullptr, *H2_ = nullptr, *Y_ = nullptr;
  float *dY_ = nullptr, *dA2_ = nullptr, *dA1_ = nullptr, *Zb_ = nullptr, *Tb_ = nullptr;
  double *scale_ = nullptr, *lossbuf_ = nullptr, *evalbuf_ = nullptr;
  int *perm_ = nullptr, *vidx_ = nullptr;
  float *Zneg_ = nullptr, *Yneg_ = nullptr;
  int best_epoch_ = 0;
  float wd_ = 0.0f;
  float *Pbest_ = nullptr;
  static constexpr int kEvalParts = 1024;
  int t_ = 0;
  bool trained_ = false;
  unsigned long long seed_ = 0;
  std::map<std::tuple<int, int, int, long, long, long, long>, GemmDesc *> descs_;
  int act_ = 0;
  bool wdirty_ = true;
  // BF16 weights [W1 | W2 | W3] and activation / gradient operands
  __bf16 *Wh_ = nullptr, *Zh_ = nullptr, *H1h_ = nullptr, *H2h_ = nullptr, *dYh_ = nullptr,
         *dA2h_ = nullptr, *dA1h_ = nullptr;
};

}  // namespace v3d

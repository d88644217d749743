// Learned error-indicator surrogate: a per-element message-passing network over the
// face-neighbour graph of the tets (no counterpart in the reference; it stands in for
// the dual solve of estimate_error, src/em_fwd.cc:1102-1180, in intermediate adaptive
// cycles).
//
//   h0 = relu(F W0 + b0)                     F: [T][16] standardized features
//   h_l = relu([h_{l-1}, mean_nb h_{l-1}] W_l + b_l),  l = 1..L
//   y = h_L wo + bo                          y ~ standardized log10 indicator
//
// GEMM operands are rounded to bf16 and multiplied on the matrix cores
// (v_mfma_f32_16x16x16_bf16, fp32 accumulate); activations, gradients and the Adam
// master weights are fp32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"

namespace emfem {

constexpr int kSurFeat = 16;   // feature width
constexpr int kSurHidden = 64;
constexpr int kSurLayers = 3;  // message-passing layers after the input layer

// Parameter layout in one flat fp32 vector.
struct SurrogateLayout {
  size_t w0, b0, wl[kSurLayers], bl[kSurLayers], wo, bo, total;
  SurrogateLayout();
};

// C[kc] (M x N) = sum_{k in chunk kc} A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn] (fp32 in memory, bf16 on the matrix cores, fp32
// accumulate).  kchunk = K for no split; partial c of a split is C + c * M * N.
// Epilogue (no split only): + bias[n], then relu when `relu`.
void gemm_bf16(hipStream_t s, int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
               int64_t sbk, int64_t sbn, float *C, int kchunk, const float *bias, bool relu);

// out[i] = sum_c parts[c][i] for c < n_parts (fixed order).
void sum_parts(hipStream_t s, int n, int n_parts, const float *parts, float *out);

// One forward/backward pass over a graph of T elements.
class SurrogateNet {
 public:
  SurrogateNet();
  void init(uint32_t seed, hipStream_t s);
  size_t n_params() const { return lay_.total; }
  void set_params(const float *host, hipStream_t s);
  void get_params(float *host, hipStream_t s) const;
  // y[T] for features F [T][16] (device) and neighbours nb [T][4] (device, -1 = none)
  void forward(hipStream_t s, int T, const float *F, const int *nb, float *y);
  // loss = mean (y - target)^2 over T; gradient into grad_ (device); returns loss
  double loss_grad(hipStream_t s, int T, const float *F, const int *nb, const float *target);
  void adam_step(hipStream_t s, double lr, int step);
  const float *grad() const { return grad_.get(); }
  float *params() { return p_.get(); }

 private:
  void ensure(int T);
  SurrogateLayout lay_;
  DevBuf<float> p_, grad_, m1_, m2_;
  int cap_ = 0;
  DevBuf<float> h_[kSurLayers + 1], mean_[kSurLayers + 1], y_, dy_, dh_, dcat_, part_;
  DevBuf<double> loss_;
};

// Per-tet features of the surrogate (standardization not applied): log10 of the
// conductivity diagonal, the three normalised off-diagonal terms, log10 frequency,
// log10 volume, log10 aspect (circumradius / (3 inradius)), log10 |E| and
// log10 |curl E| over the XY and YX solutions (interleaved columns at e_xy / e_yx with
// row stride ld), log10(1 + d/h) and log10(1 + d) of the distance d from the centroid
// to the nearest of the nrx receivers rx [nrx][3] (h = volume^(1/3)), log10 of
// sum_r h / (d_r + h) over all receivers, and log10 of the XY and YX estimator
// indicators eta_xy, eta_yx (eprim [2][nt] holds their squares): the primal factors of
// the goal-oriented indicator, which the surrogate cycles compute exactly.
void surrogate_features(hipStream_t s, int nt, const int *tets, const int *tet2edge, const double *pts,
                        const int *attr, const double *sig6, double freq, const double *e_xy, const double *e_yx,
                        int ld, const double *rx, int nrx, const double *eprim, float *F);

// F[t][j] = (F[t][j] - mean[j]) / std[j] on the device.
void standardize(hipStream_t s, int nt, const float *mean, const float *stdv, float *F);

}  // namespace emfem

// C ABI of the learned indicator surrogate (see surrogate.h and emfem_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "../../include/emfem_amd.h"
#include "context.h"

using namespace emfem;
using namespace emfem_ctx;

namespace {

constexpr int kNorm = 2 * kSurFeat + 2;

void require_sur(const emfem_context *c) {
  EMF_REQUIRE(c->sur != nullptr, kErrState, "no surrogate: call emfem_surrogate_init first");
}

// raw features of the current mesh and the XY / YX solution slots, on the device
// (the XY and YX estimator indicators, inputs of the last two features, are formed
// here too)
void device_features(emfem_context *c, int fidx, int sxy, int syx, DevBuf<float> &F) {
  require_system(c);
  require_slot(c, sxy);
  require_slot(c, syx);
  require_freq(c, fidx);
  hipStream_t s = c->stream;
  F.resize((size_t)c->T.nt * kSurFeat);
  const int ld = 2 * c->nslots;
  DevBuf<double> rx;
  if (!c->rx.empty()) rx.upload(c->rx, s);
  const int slots[2] = {sxy, syx};
  DevBuf<int> d_slots;
  d_slots.upload(slots, 2, s);
  DevBuf<double> eprim((size_t)2 * c->T.nt);
  estimate(s, c->T.nt, c->d_tets.get(), c->d_t2e.get(), c->d_nb.get(), c->d_non.get(), c->d_pts.get(),
           c->d_attr.get(), c->d_s6.get(), 2 * kPi * c->freqs[fidx], 2, d_slots.get(), c->d_X.get(), ld, eprim.get());
  surrogate_features(s, c->T.nt, c->d_tets.get(), c->d_t2e.get(), c->d_pts.get(), c->d_attr.get(), c->d_s6.get(),
                     c->freqs[fidx], c->d_X.get() + 2 * sxy, c->d_X.get() + 2 * syx, ld, rx.get(),
                     (int)(c->rx.size() / 3), eprim.get(), F.get());
  EMF_HIP(hipStreamSynchronize(s));
}

}  // namespace

namespace emfem_ctx {

// Surrogate indicator (10^(y std + mean)) for the current mesh and solutions.
std::vector<double> surrogate_indicator(emfem_context *c, int fidx, int sxy, int syx) {
  require_sur(c);
  EMF_REQUIRE(c->sur_norm.size() == (size_t)kNorm, kErrState, "the surrogate has not been trained");
  hipStream_t s = c->stream;
  DevBuf<float> F, norm, y((size_t)c->T.nt);
  device_features(c, fidx, sxy, syx, F);
  norm.upload(c->sur_norm, s);
  standardize(s, c->T.nt, norm.get(), norm.get() + kSurFeat, F.get());
  c->sur->forward(s, c->T.nt, F.get(), c->d_nb.get(), y.get());
  std::vector<float> hy(c->T.nt);
  y.download(hy.data(), hy.size(), s);
  EMF_HIP(hipStreamSynchronize(s));
  const double mu = c->sur_norm[2 * kSurFeat], sd = c->sur_norm[2 * kSurFeat + 1];
  std::vector<double> out(c->T.nt);
  for (int t = 0; t < c->T.nt; ++t) out[t] = std::pow(10.0, hy[t] * sd + mu);
  return out;
}

// Spearman rank correlation (average ranks for ties).
double rank_correlation(const std::vector<double> &a, const std::vector<double> &b) {
  const size_t n = a.size();
  auto ranks = [&](const std::vector<double> &v) {
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return v[i] < v[j]; });
    std::vector<double> r(n);
    for (size_t i = 0; i < n;) {
      size_t j = i;
      while (j + 1 < n && v[idx[j + 1]] == v[idx[i]]) ++j;
      for (size_t k = i; k <= j; ++k) r[idx[k]] = 0.5 * (i + j);
      i = j + 1;
    }
    return r;
  };
  const std::vector<double> ra = ranks(a), rb = ranks(b);
  const double m = 0.5 * (n - 1.0);
  double sab = 0.0, saa = 0.0, sbb = 0.0;
  for (size_t i = 0; i < n; ++i) {
    sab += (ra[i] - m) * (rb[i] - m);
    saa += (ra[i] - m) * (ra[i] - m);
    sbb += (rb[i] - m) * (rb[i] - m);
  }
  return saa > 0 && sbb > 0 ? sab / std::sqrt(saa * sbb) : 0.0;
}

}  // namespace emfem_ctx

extern "C" {

int emfem_surrogate_init(emfem_context *c, unsigned seed) {
  return guarded(c, [&]() {
    require_device(c);
    c->sur.reset(new SurrogateNet());
    c->sur->init(seed, c->stream);
    c->sur_norm.clear();
    c->sur_rejected = false;
  });
}

int emfem_surrogate_params(emfem_context *c, float *get, const float *set, int *n) {
  return guarded(c, [&]() {
    require_sur(c);
    const int np = (int)c->sur->n_params();
    if (n) *n = np + kNorm;
    if (set) {
      c->sur->set_params(set, c->stream);
      c->sur_norm.assign(set + np, set + np + kNorm);
      c->sur_rejected = false;
    }
    if (get) {
      c->sur->get_params(get, c->stream);
      for (int i = 0; i < kNorm; ++i) get[np + i] = i < (int)c->sur_norm.size() ? c->sur_norm[i] : 0.f;
    }
  });
}

int emfem_surrogate_features(emfem_context *c, int fidx, int slot_xy, int slot_yx, float *feats) {
  return guarded(c, [&]() {
    EMF_REQUIRE(feats != nullptr, kErrArg, "null feature array");
    DevBuf<float> F;
    device_features(c, fidx, slot_xy, slot_yx, F);
    F.download(feats, (size_t)c->T.nt * kSurFeat, c->stream);
    EMF_HIP(hipStreamSynchronize(c->stream));
  });
}

int emfem_surrogate_train(emfem_context *c, int n, const float *feats, const int *nb, const double *indicator,
                          int epochs, double lr, double *loss) {
  return guarded(c, [&]() {
    require_sur(c);
    EMF_REQUIRE(n > 0 && feats && nb && indicator && epochs >= 0, kErrArg, "bad training arguments");
    for (int i = 0; i < n * 4; ++i)
      EMF_REQUIRE(nb[i] >= -1 && nb[i] < n, kErrArg, "neighbour index out of range");
    // standardization of the features and of log10 of the indicator
    std::vector<float> norm(kNorm, 0.f);
    std::vector<double> acc(kSurFeat, 0.0), acc2(kSurFeat, 0.0);
    for (int t = 0; t < n; ++t)
      for (int j = 0; j < kSurFeat; ++j) {
        acc[j] += feats[(size_t)t * kSurFeat + j];
        acc2[j] += (double)feats[(size_t)t * kSurFeat + j] * feats[(size_t)t * kSurFeat + j];
      }
    for (int j = 0; j < kSurFeat; ++j) {
      const double m = acc[j] / n, v = std::max(0.0, acc2[j] / n - m * m);
      norm[j] = (float)m;
      norm[kSurFeat + j] = v > 1e-12 ? (float)std::sqrt(v) : 0.f;
    }
    std::vector<float> y(n);
    double ym = 0.0, y2 = 0.0;
    for (int t = 0; t < n; ++t) {
      y[t] = (float)std::log10(std::max(indicator[t], 1e-300));
      ym += y[t];
      y2 += (double)y[t] * y[t];
    }
    ym /= n;
    const double ysd = std::sqrt(std::max(1e-12, y2 / n - ym * ym));
    for (auto &v : y) v = (float)((v - ym) / ysd);
    norm[2 * kSurFeat] = (float)ym;
    norm[2 * kSurFeat + 1] = (float)ysd;
    hipStream_t s = c->stream;
    DevBuf<float> F, tg, dnorm;
    DevBuf<int> dnb;
    F.upload(feats, (size_t)n * kSurFeat, s);
    tg.upload(y, s);
    dnb.upload(nb, (size_t)n * 4, s);
    dnorm.upload(norm, s);
    standardize(s, n, dnorm.get(), dnorm.get() + kSurFeat, F.get());
    for (int e = 0; e < epochs; ++e) {
      const double l = c->sur->loss_grad(s, n, F.get(), dnb.get(), tg.get());
      c->sur->adam_step(s, lr, e + 1);
      if (loss) loss[e] = l;
    }
    EMF_HIP(hipStreamSynchronize(s));
    c->sur_norm = norm;
    c->sur_rejected = false;
  });
}

int emfem_surrogate_predict(emfem_context *c, int fidx, int slot_xy, int slot_yx, double *indicator) {
  return guarded(c, [&]() {
    EMF_REQUIRE(indicator != nullptr, kErrArg, "null indicator array");
    const std::vector<double> v = surrogate_indicator(c, fidx, slot_xy, slot_yx);
    std::copy(v.begin(), v.end(), indicator);
  });
}

int emfem_surrogate_eval(emfem_context *c, int n, const float *feats, const int *nb, const float *target, float *y,
                         double *loss, float *grad) {
  return guarded(c, [&]() {
    require_sur(c);
    EMF_REQUIRE(n > 0 && feats && nb, kErrArg, "bad evaluation arguments");
    for (int i = 0; i < n * 4; ++i)
      EMF_REQUIRE(nb[i] >= -1 && nb[i] < n, kErrArg, "neighbour index out of range");
    hipStream_t s = c->stream;
    DevBuf<float> F, tg, dy((size_t)n);
    DevBuf<int> dnb;
    F.upload(feats, (size_t)n * kSurFeat, s);
    dnb.upload(nb, (size_t)n * 4, s);
    if (target && (loss || grad)) {
      tg.upload(target, (size_t)n, s);
      const double l = c->sur->loss_grad(s, n, F.get(), dnb.get(), tg.get());
      if (loss) *loss = l;
      if (grad) {
        EMF_HIP(hipMemcpyAsync(grad, c->sur->grad(), sizeof(float) * c->sur->n_params(), hipMemcpyDeviceToHost, s));
      }
    }
    if (y) {
      c->sur->forward(s, n, F.get(), dnb.get(), dy.get());
      dy.download(y, (size_t)n, s);
    }
    EMF_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// Learned error-indicator surrogate (see surrogate.h): bf16 matrix-core GEMMs,
// neighbour means, the hand-written backward pass, Adam, and the per-tet features.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>

#include "device.h"
#include "fe.h"
#include "surrogate.h"

namespace emfem {

SurrogateLayout::SurrogateLayout() {
  size_t o = 0;
  w0 = o;
  o += (size_t)kSurFeat * kSurHidden;
  b0 = o;
  o += kSurHidden;
  for (int l = 0; l < kSurLayers; ++l) {
    wl[l] = o;
    o += (size_t)2 * kSurHidden * kSurHidden;
    bl[l] = o;
    o += kSurHidden;
  }
  wo = o;
  o += kSurHidden;
  bo = o;
  o += 1;
  total = o;
}

namespace {

constexpr int kBlock = 256;
constexpr int kCat = 2 * kSurHidden;  // [h, mean_nb h] row width

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// One wave per 16 x 16 output tile, four waves per block along N, 32 k per step
// (v_mfma_f32_16x16x32_bf16).  Lane l holds A(m0 + l % 16, k0 + 8 (l / 16) + j) and
// B(k0 + 8 (l / 16) + j, n0 + l % 16), j < 8, and receives
// D(m0 + 4 (l / 16) + j, n0 + l % 16), j < 4.
__global__ __launch_bounds__(kBlock) void k_gemm_bf16(int M, int N, int K, const float *__restrict__ A, int64_t sam,
                                                      int64_t sak, const float *__restrict__ B, int64_t sbk,
                                                      int64_t sbn, float *C, int64_t ldc, int kchunk,
                                                      const float *__restrict__ bias, int relu) {
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 64 + wave * 16;
  if (n0 >= N) return;  // whole wave
  const int kc = blockIdx.z;
  const int kb = kc * kchunk, ke = min(K, kb + kchunk);
  const int r = lane % 16, q = lane / 16;
  const int m = m0 + r, n = n0 + r;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    s16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + 8 * q + j;
      const bool kin = k < ke;
      a[j] = (kin && m < M) ? to_bf16(A[(int64_t)m * sam + (int64_t)k * sak]) : (short)0;
      b[j] = (kin && n < N) ? to_bf16(B[(int64_t)k * sbk + (int64_t)n * sbn]) : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int mm = m0 + 4 * q + j;
    if (mm >= M || n >= N) continue;
    float v = acc[j];
    if (split) {
      C[(int64_t)kc * M * N + (int64_t)mm * N + n] = v;
    } else {
      if (bias) v += bias[n];
      if (relu) v = fmaxf(v, 0.f);
      C[(int64_t)mm * ldc + n] = v;
    }
  }
}

__global__ void k_sum_parts(int n, int n_parts, const float *__restrict__ parts, float *out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < n_parts; ++c) s += parts[(size_t)c * n + i];
  out[i] = (float)s;
}

// cat[t][64 + j] = mean over the face neighbours n of t of cat[n][j] (0 without any)
__global__ void k_nb_mean(int T, const int *__restrict__ nb, float *cat) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= T * kSurHidden) return;
  const int t = i / kSurHidden, j = i % kSurHidden;
  float s = 0.f;
  int d = 0;
  for (int f = 0; f < 4; ++f) {
    const int n = nb[t * 4 + f];
    if (n >= 0) {
      s += cat[(size_t)n * kCat + j];
      ++d;
    }
  }
  cat[(size_t)t * kCat + kSurHidden + j] = d ? s / d : 0.f;
}

// dh[t][j] = (dcat[t][j] + sum_{n in nb(t)} dcat[n][64 + j] / deg(n)) * (h[t][j] > 0)
// (face adjacency is symmetric, so the transpose of the mean is a gather); the
// relu mask of the layer that produced h is applied when mask is set.
__global__ void k_cat_back(int T, const int *__restrict__ nb, const float *__restrict__ dcat,
                           const float *__restrict__ cat, int mask, float *dh) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= T * kSurHidden) return;
  const int t = i / kSurHidden, j = i % kSurHidden;
  float s = dcat[(size_t)t * kCat + j];
  for (int f = 0; f < 4; ++f) {
    const int n = nb[t * 4 + f];
    if (n < 0) continue;
    int d = 0;
    for (int g = 0; g < 4; ++g) d += nb[n * 4 + g] >= 0;
    s += dcat[(size_t)n * kCat + kSurHidden + j] / d;
  }
  if (mask && !(cat[(size_t)t * kCat + j] > 0.f)) s = 0.f;
  dh[(size_t)t * kSurHidden + j] = s;
}

// dh[t][j] = dy[t] * wo[j] * (h[t][j] > 0)
__global__ void k_out_back(int T, const float *__restrict__ dy, const float *__restrict__ wo,
                           const float *__restrict__ cat, float *dh) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= T * kSurHidden) return;
  const int t = i / kSurHidden, j = i % kSurHidden;
  dh[i] = cat[(size_t)t * kCat + j] > 0.f ? dy[t] * wo[j] : 0.f;
}

// dy = 2 (y - target) / T ; per-block partial sums of (y - target)^2 in double
__global__ __launch_bounds__(kBlock) void k_mse(int T, const float *__restrict__ y, const float *__restrict__ tg,
                                                float *dy, double *part) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < T; t += gridDim.x * kBlock) {
    const float e = y[t] - tg[t];
    dy[t] = 2.f * e / T;
    acc += (double)e * e;
  }
  sh[threadIdx.x] = acc;
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// column sums of X [T][ncol] (row stride ldx), per block of rows: part[b][c]
__global__ void k_colsum(int T, int ncol, const float *__restrict__ X, int ldx, int rows_per_block, float *part) {
  const int c = threadIdx.x;
  if (c >= ncol) return;
  const int r0 = blockIdx.x * rows_per_block, r1 = min(T, r0 + rows_per_block);
  double s = 0.0;
  for (int r = r0; r < r1; ++r) s += X[(size_t)r * ldx + c];
  part[(size_t)blockIdx.x * ncol + c] = (float)s;
}

__global__ void k_adam(int n, float *p, const float *__restrict__ g, float *m1, float *m2, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

}  // namespace

void gemm_bf16(hipStream_t s, int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
               int64_t sbk, int64_t sbn, float *C, int kchunk, const float *bias, bool relu) {
  if (M <= 0 || N <= 0) return;
  kchunk = std::max(32, (kchunk + 31) / 32 * 32);  // whole 32-deep MFMA steps
  const int nch = std::max(1, (K + kchunk - 1) / kchunk);
  dim3 grid((M + 15) / 16, (N + 63) / 64, nch);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(kBlock), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, (int64_t)N, kchunk,
                     bias, relu ? 1 : 0);
  EMF_HIP(hipGetLastError());
}

namespace {
// gemm with an explicit output row stride (no split)
void gemm_ld(hipStream_t s, int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
             int64_t sbk, int64_t sbn, float *C, int64_t ldc, const float *bias, bool relu) {
  dim3 grid((M + 15) / 16, (N + 63) / 64, 1);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(kBlock), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                     std::max(16, (K + 15) / 16 * 16), bias, relu ? 1 : 0);
  EMF_HIP(hipGetLastError());
}
}  // namespace

void sum_parts(hipStream_t s, int n, int n_parts, const float *parts, float *out) {
  hipLaunchKernelGGL(k_sum_parts, dim3(grid_for(n)), dim3(kBlock), 0, s, n, n_parts, parts, out);
  EMF_HIP(hipGetLastError());
}

SurrogateNet::SurrogateNet() {
  p_.resize(lay_.total);
  grad_.resize(lay_.total);
  m1_.resize(lay_.total);
  m2_.resize(lay_.total);
}

void SurrogateNet::init(uint32_t seed, hipStream_t s) {
  std::mt19937 g(seed);
  std::vector<float> p(lay_.total, 0.f);
  auto glorot = [&](size_t off, int fan_in, int fan_out) {
    const float a = std::sqrt(6.f / (fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < fan_in * fan_out; ++i) p[off + i] = u(g);
  };
  glorot(lay_.w0, kSurFeat, kSurHidden);
  for (int l = 0; l < kSurLayers; ++l) glorot(lay_.wl[l], 2 * kSurHidden, kSurHidden);
  glorot(lay_.wo, kSurHidden, 1);
  p_.upload(p, s);
  m1_.zero(s);
  m2_.zero(s);
  EMF_HIP(hipStreamSynchronize(s));
}

void SurrogateNet::set_params(const float *host, hipStream_t s) {
  p_.upload(host, lay_.total, s);
  m1_.zero(s);
  m2_.zero(s);
  EMF_HIP(hipStreamSynchronize(s));
}

void SurrogateNet::get_params(float *host, hipStream_t s) const {
  p_.download(host, lay_.total, s);
  EMF_HIP(hipStreamSynchronize(s));
}

void SurrogateNet::ensure(int T) {
  if (T <= cap_) return;
  cap_ = T;
  for (int l = 0; l <= kSurLayers; ++l) h_[l].resize((size_t)T * kCat);
  y_.resize(T);
  dy_.resize(T);
  dh_.resize((size_t)T * kSurHidden);
  dcat_.resize((size_t)T * kCat);
  loss_.resize(1024);
}

// h_[0] = [h0, mean h0], h_[l] = [h_l, mean h_l] (the last layer's mean unused)
void SurrogateNet::forward(hipStream_t s, int T, const float *F, const int *nb, float *y) {
  ensure(T);
  const float *P = p_.get();
  gemm_ld(s, T, kSurHidden, kSurFeat, F, kSurFeat, 1, P + lay_.w0, kSurHidden, 1, h_[0].get(), kCat, P + lay_.b0,
          true);
  for (int l = 0; l < kSurLayers; ++l) {
    hipLaunchKernelGGL(k_nb_mean, dim3(grid_for((int64_t)T * kSurHidden)), dim3(kBlock), 0, s, T, nb, h_[l].get());
    gemm_ld(s, T, kSurHidden, kCat, h_[l].get(), kCat, 1, P + lay_.wl[l], kSurHidden, 1, h_[l + 1].get(), kCat,
            P + lay_.bl[l], true);
  }
  gemm_ld(s, T, 1, kSurHidden, h_[kSurLayers].get(), kCat, 1, P + lay_.wo, 1, 0, y ? y : y_.get(), 1, P + lay_.bo,
          false);
  EMF_HIP(hipGetLastError());
}

double SurrogateNet::loss_grad(hipStream_t s, int T, const float *F, const int *nb, const float *target) {
  ensure(T);
  forward(s, T, F, nb, y_.get());
  const float *P = p_.get();
  float *G = grad_.get();
  const int nbk = std::min(1024, grid_for(T));
  hipLaunchKernelGGL(k_mse, dim3(nbk), dim3(kBlock), 0, s, T, y_.get(), target, dy_.get(), loss_.get());
  // split-K chunk over elements for the weight gradients
  const int chunk = 4096;
  const int nch = (T + chunk - 1) / chunk;
  part_.resize((size_t)nch * kCat * kSurHidden + 64);
  auto wgrad = [&](const float *X, int64_t ldx, int K, const float *D, int64_t ldd, int N, float *out) {
    // out[K][N] = X^T D over the T rows
    gemm_bf16(s, K, N, T, X, 1, ldx, D, ldd, 1, part_.get(), chunk, nullptr, false);
    sum_parts(s, K * N, nch, part_.get(), out);
  };
  auto bgrad = [&](const float *D, int64_t ldd, int N, float *out) {
    const int rpb = 512, nb2 = (T + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_colsum, dim3(nb2), dim3(64), 0, s, T, N, D, (int)ldd, rpb, part_.get());
    sum_parts(s, N, nb2, part_.get(), out);
  };
  // output layer
  wgrad(h_[kSurLayers].get(), kCat, kSurHidden, dy_.get(), 1, 1, G + lay_.wo);
  bgrad(dy_.get(), 1, 1, G + lay_.bo);
  hipLaunchKernelGGL(k_out_back, dim3(grid_for((int64_t)T * kSurHidden)), dim3(kBlock), 0, s, T, dy_.get(),
                     P + lay_.wo, h_[kSurLayers].get(), dh_.get());
  // message-passing layers: dh_ holds dZ of layer l (relu mask applied)
  for (int l = kSurLayers - 1; l >= 0; --l) {
    wgrad(h_[l].get(), kCat, kCat, dh_.get(), kSurHidden, kSurHidden, G + lay_.wl[l]);
    bgrad(dh_.get(), kSurHidden, kSurHidden, G + lay_.bl[l]);
    // dcat = dZ W_l^T : B(k, n) = W_l[n][k]
    gemm_ld(s, T, kCat, kSurHidden, dh_.get(), kSurHidden, 1, P + lay_.wl[l], 1, kSurHidden, dcat_.get(), kCat,
            nullptr, false);
    hipLaunchKernelGGL(k_cat_back, dim3(grid_for((int64_t)T * kSurHidden)), dim3(kBlock), 0, s, T, nb, dcat_.get(),
                       h_[l].get(), 1, dh_.get());
  }
  // input layer
  wgrad(F, kSurFeat, kSurFeat, dh_.get(), kSurHidden, kSurHidden, G + lay_.w0);
  bgrad(dh_.get(), kSurHidden, kSurHidden, G + lay_.b0);
  std::vector<double> parts(nbk);
  loss_.download(parts.data(), nbk, s);
  EMF_HIP(hipStreamSynchronize(s));
  double loss = 0.0;
  for (double v : parts) loss += v;
  return loss / T;
}

void SurrogateNet::adam_step(hipStream_t s, double lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(lay_.total)), dim3(kBlock), 0, s, (int)lay_.total, p_.get(), grad_.get(),
                     m1_.get(), m2_.get(), (float)lr, b1, b2, c1, c2, 1e-8f);
  EMF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------------

namespace {

__global__ void k_features(int nt, const int *__restrict__ tets, const int *__restrict__ t2e,
                           const double *__restrict__ pts, const int *__restrict__ attr, const double *__restrict__ s6,
                           double freq, const double *__restrict__ exy, const double *__restrict__ eyx, int ld,
                           const double *__restrict__ rx, int nrx, const double *__restrict__ eprim, float *F) {
  const int t = xcd_block() * kBlock + threadIdx.x;
  if (t >= nt) return;
  int vid[4];
  double x[4][3];
  for (int k = 0; k < 4; ++k) {
    vid[k] = tets[t * 4 + k];
    for (int d = 0; d < 3; ++d) x[k][d] = pts[(size_t)vid[k] * 3 + d];
  }
  TetGeom G;
  tet_geometry(x, G);
  const double vol = std::fabs(G.det) / 6.0;
  const double *sg = s6 + (size_t)attr[t] * 6;
  float *f = F + (size_t)t * kSurFeat;
  const double sx = fmax(sg[0], 1e-30), sy = fmax(sg[1], 1e-30), sz = fmax(sg[2], 1e-30);
  f[0] = (float)log10(sx);
  f[1] = (float)log10(sy);
  f[2] = (float)log10(sz);
  f[3] = (float)(sg[3] / sqrt(sx * sy));
  f[4] = (float)(sg[4] / sqrt(sx * sz));
  f[5] = (float)(sg[5] / sqrt(sy * sz));
  f[6] = (float)log10(freq);
  f[7] = (float)log10(fmax(vol, 1e-30));
  // aspect: circumradius / (3 inradius); inradius = 3 V / total face area
  double area = 0.0;
  for (int k = 0; k < 4; ++k) {
    const double gx = G.g[k][0], gy = G.g[k][1], gz = G.g[k][2];
    area += sqrt(gx * gx + gy * gy + gz * gz) * std::fabs(G.det) / 2.0;  // |grad lambda_k| = area_k / (3 V)
  }
  const double rin = 3.0 * vol / fmax(area, 1e-300);
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = x[1][d] - x[0][d];
    b[d] = x[2][d] - x[0][d];
    c[d] = x[3][d] - x[0][d];
  }
  const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2],
               cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  double bxc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  double cxa[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
  double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  double num[3];
  for (int d = 0; d < 3; ++d) num[d] = aa * bxc[d] + bb * cxa[d] + cc * axb[d];
  const double rcirc = sqrt(num[0] * num[0] + num[1] * num[1] + num[2] * num[2]) / (2.0 * std::fabs(G.det));
  f[8] = (float)log10(fmax(rcirc / (3.0 * rin), 1e-30));
  // |E| at the centroid and |curl E| (constant) over both polarisations
  double lam[4] = {0.25, 0.25, 0.25, 0.25};
  double e2 = 0.0, c2 = 0.0;
  for (int pol = 0; pol < 2; ++pol) {
    const double *e = pol ? eyx : exy;
    double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0}, cr[3] = {0, 0, 0}, ci[3] = {0, 0, 0};
    for (int i = 0; i < 6; ++i) {
      int p, q;
      edge_ends(vid, i, p, q);
      double N[3], Cn[3];
      whitney_value(G, lam, p, q, N);
      whitney_curl(G, p, q, Cn);
      const int ed = t2e[t * 6 + i];
      const double vr = e[(size_t)ed * ld], vi = e[(size_t)ed * ld + 1];
      for (int d = 0; d < 3; ++d) {
        er[d] += vr * N[d];
        ei[d] += vi * N[d];
        cr[d] += vr * Cn[d];
        ci[d] += vi * Cn[d];
      }
    }
    for (int d = 0; d < 3; ++d) {
      e2 += er[d] * er[d] + ei[d] * ei[d];
      c2 += cr[d] * cr[d] + ci[d] * ci[d];
    }
  }
  f[9] = (float)(0.5 * log10(e2 + 1e-300));
  f[10] = (float)(0.5 * log10(c2 + 1e-300));
  // distance from the centroid to the nearest receiver, in units of the tet size:
  // the dual (goal) solution the indicator weights by is concentrated there
  double cx[3] = {0, 0, 0};
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) cx[d] += 0.25 * x[k][d];
  const double h = cbrt(fmax(vol, 1e-30));
  double dmin = 1e300, dens = 0.0;
  for (int r = 0; r < nrx; ++r) {
    const double dx = rx[r * 3] - cx[0], dy = rx[r * 3 + 1] - cx[1], dz = rx[r * 3 + 2] - cx[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    dmin = fmin(dmin, d2);
    dens += h / (sqrt(d2) + h);  // superposed point sources, relative to the tet size
  }
  f[11] = nrx > 0 ? (float)log10(sqrt(dmin) / h + 1.0) : 0.f;
  f[12] = nrx > 0 ? (float)log10(sqrt(dmin) + 1.0) : 0.f;
  f[13] = nrx > 0 ? (float)log10(dens) : 0.f;
  f[14] = (float)(0.5 * log10(eprim[t] + 1e-300));
  f[15] = (float)(0.5 * log10(eprim[(size_t)nt + t] + 1e-300));
}

__global__ void k_standardize(int n, const float *__restrict__ mean, const float *__restrict__ stdv, float *F) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int j = i % kSurFeat;
  F[i] = stdv[j] > 0.f ? (F[i] - mean[j]) / stdv[j] : 0.f;
}

}  // namespace

void surrogate_features(hipStream_t s, int nt, const int *tets, const int *tet2edge, const double *pts,
                        const int *attr, const double *sig6, double freq, const double *e_xy, const double *e_yx,
                        int ld, const double *rx, int nrx, const double *eprim, float *F) {
  if (nt == 0) return;
  hipLaunchKernelGGL(k_features, dim3(grid_for(nt)), dim3(kBlock), 0, s, nt, tets, tet2edge, pts, attr, sig6, freq,
                     e_xy, e_yx, ld, rx, nrx, eprim, F);
  EMF_HIP(hipGetLastError());
}

void standardize(hipStream_t s, int nt, const float *mean, const float *stdv, float *F) {
  const int n = nt * kSurFeat;
  if (n == 0) return;
  hipLaunchKernelGGL(k_standardize, dim3(grid_for(n)), dim3(kBlock), 0, s, n, mean, stdv, F);
  EMF_HIP(hipGetLastError());
}

}  // namespace emfem

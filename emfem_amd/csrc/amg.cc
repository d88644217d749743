// Auxiliary-space Maxwell preconditioner: host set-up.  See amg.h.
#include "amg.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>

#include "common.h"

namespace emfem {

// Transpose: column counts (relaxed atomic increments, rows in parallel), their
// prefix sum, a fill through atomic per-column cursors, and a sort of each output row
// by source row, so every column lists its rows in ascending order (the order of a
// sequential transpose) whatever the threads' interleaving.  Small matrices run on
// one thread.
HostCsr csr_transpose(const HostCsr &A) {
  HostCsr T;
  T.n = A.m;
  T.m = A.n;
  T.rp.assign(A.m + 1, 0);
  T.ci.resize(A.ci.size());
  T.v.resize(A.v.size());
  const int nth = A.ci.size() < 200000 ? 1 : host_threads();
  std::vector<int> cnt(A.m, 0);
  parallel_ranges(
      A.n,
      [&](int64_t b, int64_t e, int) {
        for (int p = A.rp[b]; p < A.rp[e]; ++p) __atomic_fetch_add(&cnt[A.ci[p]], 1, __ATOMIC_RELAXED);
      },
      nth);
  for (int c = 0; c < A.m; ++c) T.rp[c + 1] = T.rp[c] + cnt[c];
  parallel_ranges(
      A.m, [&](int64_t b, int64_t e, int) { std::copy(T.rp.begin() + b, T.rp.begin() + e, cnt.begin() + b); }, nth);
  parallel_ranges(
      A.n,
      [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r)
          for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) {
            const int q = __atomic_fetch_add(&cnt[A.ci[p]], 1, __ATOMIC_RELAXED);
            T.ci[q] = (int)r;
            T.v[q] = A.v[p];
          }
      },
      nth);
  if (nth > 1) {
    // each output row sorted by source row: insertion sort on short rows, a pair sort
    // on long ones
    parallel_ranges(
        A.m,
        [&](int64_t b, int64_t e, int) {
          std::vector<std::pair<int, double>> tmp;
          for (int64_t c = b; c < e; ++c) {
            if (T.rp[c + 1] - T.rp[c] > 32) {
              tmp.clear();
              for (int i = T.rp[c]; i < T.rp[c + 1]; ++i) tmp.emplace_back(T.ci[i], T.v[i]);
              std::sort(tmp.begin(), tmp.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
              for (int i = T.rp[c]; i < T.rp[c + 1]; ++i) {
                T.ci[i] = tmp[i - T.rp[c]].first;
                T.v[i] = tmp[i - T.rp[c]].second;
              }
              continue;
            }
            for (int i = T.rp[c] + 1; i < T.rp[c + 1]; ++i) {
              const int r = T.ci[i];
              const double v = T.v[i];
              int j = i - 1;
              while (j >= T.rp[c] && T.ci[j] > r) {
                T.ci[j + 1] = T.ci[j];
                T.v[j + 1] = T.v[j];
                --j;
              }
              T.ci[j + 1] = r;
              T.v[j + 1] = v;
            }
          }
        },
        nth);
  }
  return T;
}

// Row-parallel product with a per-thread dense accumulator; columns sorted per row.
HostCsr csr_multiply(const HostCsr &A, const HostCsr &B) {
  EMF_REQUIRE(A.m == B.n, kErrArg, "csr_multiply: shape mismatch");
  HostCsr C;
  C.n = A.n;
  C.m = B.m;
  const int nth = host_threads();
  std::vector<std::vector<int>> rows_ci(A.n);
  std::vector<std::vector<double>> rows_v(A.n);
  parallel_ranges(
      A.n,
      [&](int64_t b, int64_t e, int) {
        std::vector<double> acc(B.m, 0.0);
        std::vector<int> mark(B.m, -1), cols;
        for (int64_t r = b; r < e; ++r) {
          cols.clear();
          for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) {
            const int k = A.ci[p];
            const double a = A.v[p];
            for (int q = B.rp[k]; q < B.rp[k + 1]; ++q) {
              const int c = B.ci[q];
              if (mark[c] != (int)r) {
                mark[c] = (int)r;
                acc[c] = 0.0;
                cols.push_back(c);
              }
              acc[c] += a * B.v[q];
            }
          }
          std::sort(cols.begin(), cols.end());
          rows_ci[r] = cols;
          rows_v[r].resize(cols.size());
          for (size_t i = 0; i < cols.size(); ++i) rows_v[r][i] = acc[cols[i]];
        }
      },
      nth);
  C.rp.assign(A.n + 1, 0);
  for (int r = 0; r < A.n; ++r) C.rp[r + 1] = C.rp[r] + (int)rows_ci[r].size();
  C.ci.resize(C.rp[A.n]);
  C.v.resize(C.rp[A.n]);
  parallel_ranges(A.n, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      std::copy(rows_ci[r].begin(), rows_ci[r].end(), C.ci.begin() + C.rp[r]);
      std::copy(rows_v[r].begin(), rows_v[r].end(), C.v.begin() + C.rp[r]);
    }
  });
  return C;
}

// R A P row by row without forming A P: row i of the product accumulates
// sum_k R_ik sum_j A_kj P_j. into a per-thread dense accumulator; two passes (count,
// fill) write the CSR directly.
HostCsr csr_triple(const HostCsr &R, const HostCsr &A, const HostCsr &P) {
  EMF_REQUIRE(R.m == A.n && A.m == P.n, kErrArg, "csr_triple: shape mismatch");
  HostCsr C;
  C.n = R.n;
  C.m = P.m;
  C.rp.assign(R.n + 1, 0);
  auto row = [&](int64_t i, std::vector<double> &acc, std::vector<int> &mark, std::vector<int> &cols) {
    cols.clear();
    for (int p = R.rp[i]; p < R.rp[i + 1]; ++p) {
      const int k = R.ci[p];
      const double r = R.v[p];
      for (int q = A.rp[k]; q < A.rp[k + 1]; ++q) {
        const int j = A.ci[q];
        const double ra = r * A.v[q];
        for (int s = P.rp[j]; s < P.rp[j + 1]; ++s) {
          const int c = P.ci[s];
          if (mark[c] != (int)i) {
            mark[c] = (int)i;
            acc[c] = 0.0;
            cols.push_back(c);
          }
          acc[c] += ra * P.v[s];
        }
      }
    }
  };
  // one pass: each thread keeps its rows (sorted) in its own buffers, which are then
  // copied to their offsets; both passes use the same row partition
  const int nth = host_threads();
  std::vector<std::vector<int>> tci(nth);
  std::vector<std::vector<double>> tv(nth);
  parallel_ranges(
      R.n,
      [&](int64_t b, int64_t e, int t) {
        std::vector<double> acc(P.m);
        std::vector<int> mark(P.m, -1), cols;
        for (int64_t i = b; i < e; ++i) {
          row(i, acc, mark, cols);
          std::sort(cols.begin(), cols.end());
          C.rp[i + 1] = (int)cols.size();
          for (const int c : cols) {
            tci[t].push_back(c);
            tv[t].push_back(acc[c]);
          }
        }
      },
      nth);
  for (int i = 0; i < R.n; ++i) C.rp[i + 1] += C.rp[i];
  C.ci.resize(C.rp[R.n]);
  C.v.resize(C.rp[R.n]);
  parallel_ranges(
      R.n,
      [&](int64_t b, int64_t, int t) {
        std::copy(tci[t].begin(), tci[t].end(), C.ci.begin() + C.rp[b]);
        std::copy(tv[t].begin(), tv[t].end(), C.v.begin() + C.rp[b]);
      },
      nth);
  return C;
}

namespace {

std::vector<double> diagonal(const HostCsr &A) {
  std::vector<double> d(A.n, 0.0);
  parallel_ranges(A.n, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r)
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p)
        if (A.ci[p] == r) d[r] = A.v[p];
  });
  return d;
}

std::vector<double> l1_inverse(const HostCsr &A) {
  std::vector<double> l(A.n, 0.0);
  parallel_ranges(A.n, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      double s = 0.0;
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) s += std::abs(A.v[p]);
      l[r] = s > 0.0 ? 1.0 / s : 0.0;
    }
  });
  return l;
}

// Greedy aggregation on the strength graph: roots whose strong neighbours are all
// free seed aggregates, free nodes join a neighbouring aggregate, leftovers form
// their own.
int aggregate(const HostCsr &A, const std::vector<double> &d, double theta, std::vector<int> &agg) {
  const int n = A.n;
  // strength graph: flags and row counts per nonzero in parallel, then compaction in
  // row order (also in parallel, at the prefix-summed offsets)
  std::vector<uint8_t> strong(A.ci.size());
  std::vector<int> sp(n + 1, 0), sc;
  parallel_ranges(n, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      int cnt = 0;
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) {
        const int c = A.ci[p];
        strong[p] = c != r && std::abs(A.v[p]) >= theta * std::sqrt(std::abs(d[r] * d[c]));
        cnt += strong[p];
      }
      sp[r + 1] = cnt;
    }
  });
  for (int r = 0; r < n; ++r) sp[r + 1] += sp[r];
  sc.resize(sp[n]);
  parallel_ranges(n, [&](int64_t b, int64_t e, int) {
    for (int64_t r = b; r < e; ++r) {
      int o = sp[r];
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p)
        if (strong[p]) sc[o++] = A.ci[p];
    }
  });
  agg.assign(n, -1);
  int na = 0;
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int p = sp[i]; p < sp[i + 1]; ++p)
      if (agg[sc[p]] >= 0) {
        free_nb = false;
        break;
      }
    if (!free_nb) continue;
    agg[i] = na;
    for (int p = sp[i]; p < sp[i + 1]; ++p) agg[sc[p]] = na;
    ++na;
  }
  std::vector<int> pass2(agg);
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    for (int p = sp[i]; p < sp[i + 1]; ++p)
      if (agg[sc[p]] >= 0) {
        pass2[i] = agg[sc[p]];
        break;
      }
  }
  agg.swap(pass2);
  for (int i = 0; i < n; ++i)
    if (agg[i] < 0) agg[i] = na++;
  return na;
}

double spectral_radius_dinv_a(const HostCsr &A, const std::vector<double> &d) {
  std::vector<double> x(A.n), y(A.n);
  uint64_t s = 88172645463325252ull;
  for (auto &v : x) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    v = double(s % 1000003) / 1000003.0 - 0.5;
  }
  double lam = 1.0;
  for (int it = 0; it < 15; ++it) {
    double nrm = 0.0;
    for (int r = 0; r < A.n; ++r) {
      double acc = 0.0;
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) acc += A.v[p] * x[A.ci[p]];
      y[r] = d[r] != 0.0 ? acc / d[r] : 0.0;
      nrm += y[r] * y[r];
    }
    nrm = std::sqrt(nrm);
    double xn = 0.0;
    for (double v : x) xn += v * v;
    lam = nrm / std::sqrt(xn);
    if (nrm == 0.0) break;
    for (int r = 0; r < A.n; ++r) x[r] = y[r] / nrm;
  }
  return lam;
}

// P = (I - omega D^-1 A) T with T the aggregate indicator, omega = 4 / (3 rho).
HostCsr smoothed_prolongator(const HostCsr &A, const std::vector<double> &d, const std::vector<int> &agg, int na) {
  const double rho = spectral_radius_dinv_a(A, d);
  const double omega = 4.0 / (3.0 * std::max(rho, 1e-300));
  HostCsr P;
  P.n = A.n;
  P.m = na;
  std::vector<std::vector<int>> rc(A.n);
  std::vector<std::vector<double>> rv(A.n);
  parallel_ranges(A.n, [&](int64_t b, int64_t e, int) {
    std::vector<std::pair<int, double>> ent;
    for (int64_t r = b; r < e; ++r) {
      ent.clear();
      ent.emplace_back(agg[r], 1.0);
      const double s = d[r] != 0.0 ? -omega / d[r] : 0.0;
      for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) ent.emplace_back(agg[A.ci[p]], s * A.v[p]);
      std::sort(ent.begin(), ent.end(), [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
        return x.first < y.first;
      });
      for (size_t i = 0; i < ent.size();) {
        size_t j = i;
        double v = 0.0;
        while (j < ent.size() && ent[j].first == ent[i].first) v += ent[j++].second;
        if (v != 0.0) {
          rc[r].push_back(ent[i].first);
          rv[r].push_back(v);
        }
        i = j;
      }
    }
  });
  P.rp.assign(A.n + 1, 0);
  for (int r = 0; r < A.n; ++r) P.rp[r + 1] = P.rp[r] + (int)rc[r].size();
  P.ci.resize(P.rp[A.n]);
  P.v.resize(P.rp[A.n]);
  for (int r = 0; r < A.n; ++r) {
    std::copy(rc[r].begin(), rc[r].end(), P.ci.begin() + P.rp[r]);
    std::copy(rv[r].begin(), rv[r].end(), P.v.begin() + P.rp[r]);
  }
  return P;
}

// Runs f(0) .. f(count - 1) on the host threads, each task taken by the next free
// thread; the tasks write disjoint outputs, so results do not depend on the order.
template <typename F>
void run_tasks(int count, F &&f) {
  const int nth = std::min(count, host_threads());
  if (nth <= 1) {
    for (int k = 0; k < count; ++k) f(k);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t)
    th.emplace_back([&]() {
      for (int k = next++; k < count; k = next++) f(k);
    });
  for (auto &x : th) x.join();
}

// Dense SPD inverse by Cholesky; a vanishing pivot (singular coarse problem) is
// regularised, which only affects the null-space component.  The matrix is inverted
// per connected component of its graph (the blocks of a block-diagonal matrix, e.g. the
// auxiliary blocks and the models of a stacked batch): the Cholesky of the whole
// matrix would do the same operations on each block plus exact zeros, and the
// regularisation threshold stays relative to the largest diagonal entry of the whole
// matrix.  The components are factored in parallel, then their columns are solved in
// parallel chunks.  inv_t receives the inverse transposed (row j holds column j of the
// inverse, zero outside its component), span[2 i] and span[2 i + 1] the first and one
// past the last row of row i's component.
void dense_inverse(const HostCsr &A, std::vector<double> &inv_t, std::vector<int> &span) {
  const int n = A.n;
  double dmax = 0.0;
  std::vector<int> root(n);
  for (int i = 0; i < n; ++i) root[i] = i;
  auto find = [&](int i) {
    while (root[i] != i) i = root[i] = root[root[i]];
    return i;
  };
  for (int r = 0; r < n; ++r)
    for (int p = A.rp[r]; p < A.rp[r + 1]; ++p) {
      if (A.ci[p] == r) dmax = std::max(dmax, std::abs(A.v[p]));
      const int x = find(r), y = find(A.ci[p]);
      if (x != y) root[std::max(x, y)] = std::min(x, y);
    }
  // components, each listing its rows in ascending order
  std::vector<int> comp_of(n, -1);
  std::vector<std::vector<int>> comps;
  for (int i = 0; i < n; ++i) {
    const int r = find(i);
    if (comp_of[r] < 0) {
      comp_of[r] = (int)comps.size();
      comps.emplace_back();
    }
    comps[comp_of[r]].push_back(i);
  }
  span.assign(2 * (size_t)n, 0);
  for (const auto &idx : comps)
    for (int i : idx) {
      span[2 * (size_t)i] = idx.front();
      span[2 * (size_t)i + 1] = idx.back() + 1;
    }
  inv_t.assign((size_t)n * n, 0.0);
  // Cholesky factor L of each component and its transpose (the back substitution
  // reads L by columns)
  const int nc = (int)comps.size();
  std::vector<std::vector<double>> Ls(nc), LTs(nc);
  run_tasks(nc, [&](int kc) {
    const std::vector<int> &idx = comps[kc];
    const int m = (int)idx.size();
    std::unordered_map<int, int> loc;
    loc.reserve(2 * (size_t)m);
    for (int i = 0; i < m; ++i) loc[idx[i]] = i;
    std::vector<double> a((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i)
      for (int p = A.rp[idx[i]]; p < A.rp[idx[i] + 1]; ++p) a[(size_t)i * m + loc.at(A.ci[p])] = A.v[p];
    // symmetrise (Galerkin products are symmetric up to rounding)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < i; ++j) {
        double v = 0.5 * (a[(size_t)i * m + j] + a[(size_t)j * m + i]);
        a[(size_t)i * m + j] = a[(size_t)j * m + i] = v;
      }
    std::vector<double> &L = Ls[kc];
    L.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
      const double *Lj = L.data() + (size_t)j * m;
      double s = a[(size_t)j * m + j];
      for (int k = 0; k < j; ++k) s -= Lj[k] * Lj[k];
      if (s <= 1e-14 * dmax) s = std::max(1e-14 * dmax, 1e-300);
      const double ljj = std::sqrt(s);
      L[(size_t)j * m + j] = ljj;
      // four rows at a time: independent chains, each summed in the order of k
      int i = j + 1;
      for (; i + 4 <= m; i += 4) {
        const double *L0 = L.data() + (size_t)i * m, *L1 = L0 + m, *L2 = L1 + m, *L3 = L2 + m;
        double t0 = a[(size_t)i * m + j], t1 = a[(size_t)(i + 1) * m + j], t2 = a[(size_t)(i + 2) * m + j],
               t3 = a[(size_t)(i + 3) * m + j];
        for (int k = 0; k < j; ++k) {
          const double l = Lj[k];
          t0 -= L0[k] * l;
          t1 -= L1[k] * l;
          t2 -= L2[k] * l;
          t3 -= L3[k] * l;
        }
        L[(size_t)i * m + j] = t0 / ljj;
        L[(size_t)(i + 1) * m + j] = t1 / ljj;
        L[(size_t)(i + 2) * m + j] = t2 / ljj;
        L[(size_t)(i + 3) * m + j] = t3 / ljj;
      }
      for (; i < m; ++i) {
        const double *Li = L.data() + (size_t)i * m;
        double t = a[(size_t)i * m + j];
        for (int k = 0; k < j; ++k) t -= Li[k] * Lj[k];
        L[(size_t)i * m + j] = t / ljj;
      }
    }
    std::vector<double> &LT = LTs[kc];
    LT.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k <= i; ++k) LT[(size_t)k * m + i] = L[(size_t)i * m + k];
  });
  // kCols columns c0 + [0, kCols) of a component's inverse at a time: L y = e_c (y is
  // zero above c), then L^T x = y, the columns as independent chains over one pass of
  // L (each summed in the order of k; the rows above a column's c add exact zeros)
  constexpr int kCols = 8;
  std::vector<std::pair<int, int>> tasks;
  for (int kc = 0; kc < nc; ++kc)
    for (int c0 = 0; c0 < (int)comps[kc].size(); c0 += kCols) tasks.emplace_back(kc, c0);
  run_tasks((int)tasks.size(), [&](int task) {
    const int kc = tasks[task].first, c0 = tasks[task].second;
    const std::vector<int> &idx = comps[kc];
    const std::vector<double> &L = Ls[kc], &LT = LTs[kc];
    const int m = (int)idx.size();
    const int nk = std::min(kCols, m - c0);
    std::vector<double> y((size_t)m * kCols, 0.0);  // [row][column of the chunk]
    for (int i = c0; i < m; ++i) {
      const double *Li = L.data() + (size_t)i * m;
      double t[kCols];
      for (int q = 0; q < kCols; ++q) t[q] = (i == c0 + q) ? 1.0 : 0.0;
      for (int k = c0; k < i; ++k) {
        const double l = Li[k];
        const double *yk = y.data() + (size_t)k * kCols;
        for (int q = 0; q < kCols; ++q) t[q] -= l * yk[q];
      }
      for (int q = 0; q < kCols; ++q) y[(size_t)i * kCols + q] = t[q] / Li[i];
    }
    for (int i = m - 1; i >= 0; --i) {
      const double *LTi = LT.data() + (size_t)i * m;
      double t[kCols];
      for (int q = 0; q < kCols; ++q) t[q] = y[(size_t)i * kCols + q];
      for (int k = i + 1; k < m; ++k) {
        const double l = LTi[k];
        const double *yk = y.data() + (size_t)k * kCols;
        for (int q = 0; q < kCols; ++q) t[q] -= l * yk[q];
      }
      for (int q = 0; q < kCols; ++q) y[(size_t)i * kCols + q] = t[q] / LTi[i];
    }
    for (int q = 0; q < nk; ++q) {
      double *dst = inv_t.data() + (size_t)idx[c0 + q] * n;
      for (int i = 0; i < m; ++i) dst[idx[i]] = y[(size_t)i * kCols + q];
    }
  });
}

}  // namespace

void build_amg(HostCsr A, const AmgParams &prm, AmgHost &H) {
  H.levels.clear();
  auto tick = [&](const char *k, double t0) {
    if (prm.timing) (*prm.timing)[k] += now_ms() - t0;
  };
  while (A.n > prm.max_coarse && (int)H.levels.size() < prm.max_levels) {
    double t0 = now_ms();
    std::vector<double> d = diagonal(A);
    std::vector<int> agg;
    int na = aggregate(A, d, prm.theta, agg);
    if (na >= A.n * 0.9) na = aggregate(A, d, 0.0, agg);  // weak graph: coarsen on all couplings
    tick("amg_aggregate_ms", t0);
    if (na >= A.n * 0.9 || na == 0) break;
    AmgLevelHost L;
    t0 = now_ms();
    if ((int)H.levels.size() < prm.smooth_levels) {
      L.P = smoothed_prolongator(A, d, agg, na);
    } else {
      // plain aggregation: one unit entry per row keeps every coarse stencil as sparse
      // as the fine one (smoothed prolongators grew 14 -> 350+ nonzeros per row here
      // without changing the auxiliary-space iteration count)
      L.P.n = A.n;
      L.P.m = na;
      L.P.rp.resize(A.n + 1);
      L.P.ci.resize(A.n);
      L.P.v.assign(A.n, 1.0);
      for (int i = 0; i <= A.n; ++i) L.P.rp[i] = i;
      for (int i = 0; i < A.n; ++i) L.P.ci[i] = agg[i];
    }
    tick("amg_prolongator_ms", t0);
    t0 = now_ms();
    L.R = csr_transpose(L.P);
    L.l1inv = l1_inverse(A);
    tick("amg_transpose_ms", t0);
    t0 = now_ms();
    HostCsr Ac = csr_triple(L.R, A, L.P);
    tick("amg_rap_ms", t0);
    L.A = std::move(A);
    H.levels.push_back(std::move(L));
    A = std::move(Ac);
  }
  EMF_REQUIRE(A.n <= 4096, kErrUser,
              "auxiliary AMG stopped coarsening at " + std::to_string(A.n) + " unknowns");
  H.ncoarse = A.n;
  const double t0 = now_ms();
  dense_inverse(A, H.coarse_inv_t, H.coarse_span);
  tick("amg_coarse_ms", t0);
}

std::shared_ptr<const AmsAux> build_ams_aux(const Topology &T, bool additive) {
  EMF_REQUIRE(T.nt > 0 && T.tets.size() == (size_t)T.nt * 4 && T.pts.size() == (size_t)T.nv * 3 &&
                  T.edgelist.size() == (size_t)T.ne * 2 && T.vertex_bdr.size() == (size_t)T.nv,
              kErrState, "auxiliary spaces: incomplete topology");
  auto X = std::make_shared<AmsAux>();
  // interior vertex numbering
  std::vector<int> vmap(T.nv, -1);
  int nv0 = 0;
  for (int v = 0; v < T.nv; ++v)
    if (!T.vertex_bdr[v]) vmap[v] = nv0++;
  X->nv0 = nv0;
  EMF_REQUIRE(nv0 > 0, kErrUser, "mesh has no interior vertex");
  // the row pointers G0 and the three Pi_d share: one entry per interior endpoint
  HostVec<int> erp(T.ne + 1, 0);
  parallel_ranges(T.ne, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) erp[i + 1] = (vmap[T.edgelist[i * 2]] >= 0) + (vmap[T.edgelist[i * 2 + 1]] >= 0);
  });
  for (int e = 0; e < T.ne; ++e) erp[e + 1] += erp[e];
  auto make = [&](int comp) {
    HostCsr G;
    G.n = T.ne;
    G.m = nv0;
    G.rp = erp;
    G.ci.resize(G.rp[T.ne]);
    G.v.resize(G.rp[T.ne]);
    parallel_ranges(T.ne, [&](int64_t eb, int64_t ee, int) {
      for (int64_t e = eb; e < ee; ++e) {
        const int a = T.edgelist[e * 2], b = T.edgelist[e * 2 + 1];
        double va, vb;
        if (comp < 0) {
          va = -1.0;  // G row e: -1 at the lower vertex, +1 at the higher (src/em_la.cc:264-267)
          vb = 1.0;
        } else {
          const double t = T.pts[(size_t)b * 3 + comp] - T.pts[(size_t)a * 3 + comp];
          va = vb = 0.5 * t;
        }
        int q = G.rp[e];
        // columns ascending within the row
        const bool swap = vmap[a] >= 0 && vmap[b] >= 0 && vmap[a] > vmap[b];
        const int c0 = swap ? b : a, c1 = swap ? a : b;
        const double v0 = swap ? vb : va, v1 = swap ? va : vb;
        if (vmap[c0] >= 0) {
          G.ci[q] = vmap[c0];
          G.v[q++] = v0;
        }
        if (vmap[c1] >= 0) {
          G.ci[q] = vmap[c1];
          G.v[q++] = v1;
        }
      }
    });
    return G;
  };
  X->G0 = make(-1);
  X->G0T = csr_transpose(X->G0);
  for (int d = 0; d < 3; ++d) {
    X->P[d] = make(d);
    X->PT[d] = csr_transpose(X->P[d]);
  }
  X->additive = additive;
  const int nb = additive ? 4 : 3;
  X->nblk = nb;
  // Galerkin pattern: interior vertices u, v of a common tet (B couples the edges of a
  // tet, and P maps an edge to its interior endpoints).  Every two vertices of a tet
  // are the ends of one of its edges, so row u holds u and the other interior ends of
  // the edges at u (the rows of G0^T), built per row in parallel.
  {
    HostCsr &A = X->Apat;
    A.n = A.m = nv0;
    std::vector<int> ivmap(nv0);
    for (int v = 0; v < T.nv; ++v)
      if (vmap[v] >= 0) ivmap[vmap[v]] = v;
    const HostCsr &GT = X->G0T;
    auto other = [&](int u, int e) {
      const int a = T.edgelist[(size_t)e * 2], b = T.edgelist[(size_t)e * 2 + 1];
      return vmap[a == ivmap[u] ? b : a];
    };
    A.rp.assign(nv0 + 1, 0);
    parallel_ranges(nv0, [&](int64_t b, int64_t e, int) {
      for (int64_t u = b; u < e; ++u) {
        int k = 1;
        for (int p = GT.rp[u]; p < GT.rp[u + 1]; ++p) k += other((int)u, GT.ci[p]) >= 0;
        A.rp[u + 1] = k;
      }
    });
    for (int u = 0; u < nv0; ++u) A.rp[u + 1] += A.rp[u];
    A.ci.resize(A.rp[nv0]);
    parallel_ranges(nv0, [&](int64_t b, int64_t e, int) {
      for (int64_t u = b; u < e; ++u) {
        int q = A.rp[u];
        A.ci[q++] = (int)u;
        for (int p = GT.rp[u]; p < GT.rp[u + 1]; ++p) {
          const int v = other((int)u, GT.ci[p]);
          if (v >= 0) A.ci[q++] = v;
        }
        std::sort(A.ci.begin() + A.rp[u], A.ci.begin() + A.rp[u + 1]);
      }
    });
    // block-diagonal stack of the nblk Galerkin blocks
    HostCsr &S = X->APpat;
    const int64_t z = A.rp[nv0];
    S.n = S.m = nb * nv0;
    S.rp.resize((size_t)nb * nv0 + 1);
    S.ci.resize((size_t)nb * z);
    S.rp[0] = 0;
    for (int d = 0; d < nb; ++d) {
      parallel_ranges(nv0, [&](int64_t b, int64_t e, int) {
        for (int64_t u = b; u < e; ++u) S.rp[(size_t)d * nv0 + u + 1] = (int)(d * z + A.rp[u + 1]);
      });
      parallel_ranges(z, [&](int64_t b, int64_t e, int) {
        for (int64_t k = b; k < e; ++k) S.ci[d * z + k] = A.ci[k] + d * nv0;
      });
    }
  }
  return X;
}

void build_pi_csr(const AmsAux &X, HostCsr &Pi, HostCsr &PiT) {
  std::vector<const HostCsr *> pb;
  if (X.additive) pb.push_back(&X.G0);
  for (int d = 0; d < 3; ++d) pb.push_back(&X.P[d]);
  const int nb = (int)pb.size(), ne = X.G0.n, nv0 = X.nv0;
  Pi = HostCsr();
  Pi.n = ne;
  Pi.m = nb * nv0;
  Pi.rp.resize(ne + 1);
  // every block has G0's row structure
  parallel_ranges(ne + 1, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i) Pi.rp[i] = nb * X.G0.rp[i];
  });
  Pi.ci.resize(Pi.rp[ne]);
  Pi.v.resize(Pi.rp[ne]);
  parallel_ranges(ne, [&](int64_t eb, int64_t ee, int) {
    for (int64_t e = eb; e < ee; ++e) {
      int q = Pi.rp[e];
      for (int d = 0; d < nb; ++d)
        for (int k = pb[d]->rp[e]; k < pb[d]->rp[e + 1]; ++k) {
          Pi.ci[q] = d * nv0 + pb[d]->ci[k];
          Pi.v[q++] = pb[d]->v[k];
        }
    }
  });
  PiT = csr_transpose(Pi);
}

// The four Galerkin products P_k^T B P_k (k = G0, Pi_x, Pi_y, Pi_z) in one pass over
// B: they share the pattern Apat, so each row u accumulates the four value sets at
// positions found through a dense column map, into vals[k] (Apat layout; each thread
// zeroes its own rows first, so the pages are first touched in parallel).  Each value
// is summed in the order of csr_triple (rows of P_k^T, then of B, then of P_k).
void galerkin_blocks(const AmsAux &X, const HostCsr &B, double *const vals[4]) {
  const HostCsr &A = X.Apat;
  const HostCsr *R[4] = {&X.G0T, &X.PT[0], &X.PT[1], &X.PT[2]};
  const HostCsr *P[4] = {&X.G0, &X.P[0], &X.P[1], &X.P[2]};
  EMF_REQUIRE(B.n == X.G0.n && A.n == X.nv0, kErrArg, "Galerkin blocks: B does not match the auxiliary spaces");
  std::atomic<int64_t> missing{0};
  parallel_ranges(A.n, [&](int64_t b, int64_t e, int) {
    std::vector<int> pos(A.m, -1);
    for (int k = 0; k < 4; ++k) std::fill(vals[k] + A.rp[b], vals[k] + A.rp[e], 0.0);
    for (int64_t u = b; u < e; ++u) {
      const int r0 = A.rp[u];
      for (int t = r0; t < A.rp[u + 1]; ++t) pos[A.ci[t]] = t;
      for (int p = R[0]->rp[u]; p < R[0]->rp[u + 1]; ++p) {
        const int ed = R[0]->ci[p];
        double r[4];
        for (int k = 0; k < 4; ++k) r[k] = R[k]->v[p];
        for (int q = B.rp[ed]; q < B.rp[ed + 1]; ++q) {
          const int f = B.ci[q];
          const double bv = B.v[q];
          double ra[4];
          for (int k = 0; k < 4; ++k) ra[k] = r[k] * bv;
          for (int s2 = P[0]->rp[f]; s2 < P[0]->rp[f + 1]; ++s2) {
            const int t = pos[P[0]->ci[s2]];
            if (t < r0 || t >= A.rp[u + 1]) {  // B couples edges outside a common tet
              ++missing;
              continue;
            }
            for (int k = 0; k < 4; ++k) vals[k][t] += ra[k] * P[k]->v[s2];
          }
        }
      }
    }
  });
  EMF_REQUIRE(missing.load() == 0, kErrArg, "Galerkin blocks: B has couplings outside the tet pattern");
}

void galerkin_blocks(const AmsAux &X, const HostCsr &B, std::vector<double> (&vals)[4]) {
  const int64_t z = X.Apat.rp[X.Apat.n];
  double *out[4];
  for (int k = 0; k < 4; ++k) {
    vals[k].resize(z);
    out[k] = vals[k].data();
  }
  galerkin_blocks(X, B, out);
}

void build_ams(const Topology &T, const HostCsr &B, const AmgParams &prm_in, AmsHost &H,
               std::shared_ptr<const AmsAux> aux) {
  build_ams(
      T, [&](const AmsAux &X, double *const out[4]) { galerkin_blocks(X, B, out); }, prm_in, H, std::move(aux));
  H.l1inv = l1_inverse(B);
}

void build_ams(const Topology &T, const GalerkinFill &fill, const AmgParams &prm_in, AmsHost &H,
               std::shared_ptr<const AmsAux> aux) {
  AmgParams prm = prm_in;
  prm.timing = &H.timing;
  double t0 = now_ms();
  if (!aux || aux->additive != prm.additive || aux->G0.n != T.ne) aux = build_ams_aux(T, prm.additive);
  H.aux = aux;
  const int nv0 = aux->nv0;
  H.timing["aux_build_ms"] += now_ms() - t0;
  // Galerkin products of the auxiliary blocks (one fused pass, shared pattern), written
  // straight into the block-diagonal stack; a recycled stack from an earlier set-up on
  // this mesh keeps its pages
  t0 = now_ms();
  (void)nv0;
  const bool additive = aux->additive;
  const int64_t z = aux->Apat.rp[aux->Apat.n];
  HostCsr AP = std::move(H.recycle);
  const bool same = H.recycle_aux == aux && AP.n == aux->APpat.n && AP.m == aux->APpat.m &&
                    AP.rp.size() == aux->APpat.rp.size() && AP.ci.size() == aux->APpat.ci.size();
  H.recycle_aux.reset();
  if (!same) {
    AP.n = aux->APpat.n;
    AP.m = aux->APpat.m;
    AP.rp.assign(aux->APpat.rp.begin(), aux->APpat.rp.end());
    AP.ci.assign(aux->APpat.ci.begin(), aux->APpat.ci.end());
  }
  AP.v.resize(AP.ci.size());
  HostVec<double> vg;
  double *out[4];
  for (int k = 0; k < 4; ++k) {
    const int d = additive ? k : k - 1;
    if (d >= 0) out[k] = AP.v.data() + d * z;
  }
  if (!additive) {
    vg.resize(z);
    out[0] = vg.data();
  }
  fill(*aux, out);
  H.timing["galerkin_ms"] += now_ms() - t0;
  if (!additive) {
    t0 = now_ms();
    HostCsr AG = aux->Apat;
    AG.v = std::move(vg);
    build_amg(std::move(AG), prm, H.amgG);
    H.timing["amg_ms"] += now_ms() - t0;
  }
  t0 = now_ms();
  build_amg(std::move(AP), prm, H.amgPi);
  H.timing["amg_ms"] += now_ms() - t0;
}

}  // namespace emfem

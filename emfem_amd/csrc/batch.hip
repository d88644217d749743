// Multi-model batches: independent models solved as one block-diagonal system.
//
// The reference solves its models (and its frequency/transmitter tasks) one after
// another, or in separate MPI task groups (src/em_ctx.cc:17-43,
// src/em_fwd.cc:1273-1318).  Here the assembled systems of many contexts are stacked
// block-diagonally on one device: every SpMM, smoother sweep and AMG V-cycle runs
// once over all models (one launch per operation for the whole batch, each model at
// its own size), while every inner product, Arnoldi process, FCG recurrence and
// stopping test stays per model (segmented reductions, kernels.h Seg).  The result
// of each model is its own FGMRES solve to its own tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/emfem_amd.h"
#include "context.h"

using namespace emfem;
using namespace emfem_ctx;

namespace {

constexpr int kChunkRows = 8192;  // rows per segmented-reduction chunk

__global__ void k_fill_segments(int rows, const int *__restrict__ row0, int nseg, int *rseg) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= rows) return;
  int lo = 0, hi = nseg - 1;  // last segment whose first row <= r
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (row0[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  rseg[r] = lo;
}

}  // namespace

namespace {

// Models left in a stage below which the stage runs to completion.
constexpr int kMinStageModels = 8;

struct StageOut {
  SolveStats st;
  double setup_ms = 0.0, solve_ms = 0.0, solve_start_ms = 0.0;
  std::map<std::string, double> timing;  // set-up phases (ms)
};

// One stage: the models `S` stacked block-diagonally, warm-started from their solution
// slots when `warm`, solved until at most stop_remaining right-hand sides are left (0:
// to completion); every model's iterate is written back to its solution slots.
void solve_stage(emfem_context **ctxs, const std::vector<int> &S, int first, int n, int max_it, double rtol, bool warm,
                 int stop_remaining, double t_begin, StageOut &out) {
  emfem_context *c0 = ctxs[0];
  const int ns = (int)S.size();
  int64_t ne = 0, nnz = 0, nv = 0;
  std::vector<int> erow0(ns + 1, 0), nzoff(ns + 1, 0), voff(ns + 1, 0);
  for (int k = 0; k < ns; ++k) {
    emfem_context *c = ctxs[S[k]];
    ne += c->T.ne;
    nnz += c->nnz;
    nv += c->T.nv;
    EMF_REQUIRE(ne < INT32_MAX && nnz < INT32_MAX, kErrArg, "batch exceeds int32 indexing");
    erow0[k + 1] = (int)ne;
    nzoff[k + 1] = (int)nnz;
    voff[k + 1] = (int)nv;
  }
  hipStream_t s = c0->stream;
  const double t0 = now_ms();
  // stacked CSR (device) and B = C + M with the union topology (host) for the set-up
  DevBuf<int> rp((size_t)ne + 1), ci((size_t)nnz);
  DevBuf<double2> K((size_t)nnz);
  HostCsr B;
  B.n = B.m = (int)ne;
  B.rp.assign((size_t)ne + 1, 0);
  B.ci.resize((size_t)nnz);
  B.v.resize((size_t)nnz);
  Topology U;
  U.nv = (int)nv;
  U.ne = (int)ne;
  std::vector<int64_t> toff(ns + 1, 0);
  for (int k = 0; k < ns; ++k) toff[k + 1] = toff[k] + ctxs[S[k]]->T.nt;
  U.nt = (int)toff[ns];
  U.pts.resize((size_t)nv * 3);
  U.edgelist.resize((size_t)ne * 2);
  U.vertex_bdr.resize((size_t)nv);
  U.tets.resize((size_t)toff[ns] * 4);
  // every model fills its own ranges of the stacked arrays; the models are spread over
  // host threads, each downloading on its model's stream
  auto stack_model = [&](int k) {
    emfem_context *c = ctxs[S[k]];
    const int e0 = erow0[k], z0 = nzoff[k], v0 = voff[k];
    std::vector<int> hrp(c->T.ne + 1);
    // B = C + M formed on the device: half the bytes of (C, M) travel
    DevBuf<double> bm(c->nnz);
    make_b<double>(c->stream, c->nnz, c->d_K.get(), bm.get());
    c->d_rp.download(hrp.data(), hrp.size(), c->stream);
    c->d_ci.download(&B.ci[z0], c->nnz, c->stream);
    bm.download(&B.v[z0], c->nnz, c->stream);
    EMF_HIP(hipStreamSynchronize(c->stream));
    for (int r = 0; r < c->T.ne; ++r) B.rp[e0 + r + 1] = z0 + hrp[r + 1];
    for (int i = 0; i < c->nnz; ++i) B.ci[z0 + i] += e0;
    std::copy(c->T.pts.begin(), c->T.pts.end(), U.pts.begin() + (size_t)v0 * 3);
    std::copy(c->T.vertex_bdr.begin(), c->T.vertex_bdr.end(), U.vertex_bdr.begin() + v0);
    for (size_t i = 0; i < c->T.edgelist.size(); ++i) U.edgelist[(size_t)e0 * 2 + i] = c->T.edgelist[i] + v0;
    // tets (the pattern of the auxiliary Galerkin products)
    for (size_t i = 0; i < c->T.tets.size(); ++i) U.tets[(size_t)toff[k] * 4 + i] = c->T.tets[i] + v0;
  };
  {
    const int nth = std::max(1, std::min(ns, host_threads()));
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(nth);
    for (int t = 0; t < nth; ++t)
      th.emplace_back([&, t]() {
        try {
          for (int k = t; k < ns; k += nth) stack_model(k);
        } catch (...) {
          err[t] = std::current_exception();
        }
      });
    for (auto &x : th) x.join();
    for (auto &e : err)
      if (e) std::rethrow_exception(e);
  }
  for (int k = 0; k < ns; ++k) {
    emfem_context *c = ctxs[S[k]];
    EMF_HIP(hipMemcpyAsync(K.get() + nzoff[k], c->d_K.get(), sizeof(double2) * c->nnz, hipMemcpyDeviceToDevice, s));
  }
  rp.upload(B.rp, s);
  ci.upload(B.ci, s);
  out.timing["stack"] += now_ms() - t0;
  // segments: model k owns rows [erow0[k], erow0[k+1]); reduction chunks of at most
  // kChunkRows rows never straddle two models
  std::vector<int> chunk_row, seg_chunk(ns + 1, 0);
  for (int k = 0; k < ns; ++k) {
    seg_chunk[k] = (int)chunk_row.size();
    const int rows = erow0[k + 1] - erow0[k];
    const int nch = std::max(1, (rows + kChunkRows - 1) / kChunkRows);
    for (int j = 0; j < nch; ++j) chunk_row.push_back(erow0[k] + (int)((int64_t)rows * j / nch));
  }
  seg_chunk[ns] = (int)chunk_row.size();
  chunk_row.push_back((int)ne);
  DevBuf<int> d_row0, d_rseg((size_t)ne), d_chunk_row, d_seg_chunk;
  d_row0.upload(erow0, s);
  d_chunk_row.upload(chunk_row, s);
  d_seg_chunk.upload(seg_chunk, s);
  hipLaunchKernelGGL(k_fill_segments, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, (int)ne, d_row0.get(), ns,
                     d_rseg.get());
  EMF_HIP(hipGetLastError());
  Seg sg;
  sg.nseg = ns;
  sg.nchunks = (int)chunk_row.size() - 1;
  sg.rseg = d_rseg.get();
  sg.chunk_row = d_chunk_row.get();
  sg.seg_chunk = d_seg_chunk.get();
  // auxiliary-space set-up on the union (aggregates never cross models: no
  // couplings between blocks) and the inner solver
  PcOptions o = pc_options(c0);
  AmsHost H;
  build_ams(U, B, o.amg, H);
  for (auto &kv : H.timing) out.timing[kv.first.substr(0, kv.first.size() - 3)] += kv.second;
  const double tp = now_ms();
  auto pc = make_inner_pc(c0->opt("inner_precision") == "fp32", H, (int)ne, (int)nnz, rp.get(), ci.get(), K.get(), o,
                          s);
  out.timing["device_pc"] += now_ms() - tp;
  pc->set_use_graph(c0->opt("use_graph") == "true");
  // gather the right-hand sides (and the warm-start iterates), solve, scatter
  const int nc = 2 * n;
  DevBuf<double> b((size_t)ne * nc), x((size_t)ne * nc);
  for (int k = 0; k < ns; ++k) {
    emfem_context *c = ctxs[S[k]];
    axpby<double, double>(s, c->T.ne, nc, 1.0, c->d_S.get() + 2 * first, 2 * c->nslots, 0.0,
                          b.get() + (size_t)erow0[k] * nc, nc);
    if (warm)
      axpby<double, double>(s, c->T.ne, nc, 1.0, c->d_X.get() + 2 * first, 2 * c->nslots, 0.0,
                            x.get() + (size_t)erow0[k] * nc, nc);
  }
  std::vector<double> tol((size_t)ns * n, rtol);
  OuterOptions oo;
  oo.restart = c0->opti("A_ksp_gmres_restart");
  oo.max_it = max_it;
  oo.fp64_fallback = c0->opt("fp64_fallback") == "true";
  oo.presb = c0->opt("A_pc_type") == "presb";
  oo.floor_factor = c0->optd("true_residual_slack");
  oo.x_given = warm;
  oo.stop_remaining = stop_remaining;
  EMF_HIP(hipStreamSynchronize(s));
  const double t1 = now_ms();
  fgmres(s, (int)ne, (int)nnz, rp.get(), ci.get(), K.get(), *pc, n, b.get(), x.get(), tol, oo, out.st, sg);
  for (int k = 0; k < ns; ++k) {
    emfem_context *c = ctxs[S[k]];
    axpby<double, double>(s, c->T.ne, nc, 1.0, x.get() + (size_t)erow0[k] * nc, nc, 0.0, c->d_X.get() + 2 * first,
                          2 * c->nslots);
  }
  EMF_HIP(hipStreamSynchronize(s));
  const double t2 = now_ms();
  out.setup_ms = t1 - t0;
  out.solve_ms = t2 - t1;
  out.solve_start_ms = t1 - t_begin;
}

}  // namespace

extern "C" int emfem_solve_batch(emfem_context **ctxs, int n_ctx, int first, int n, int max_it, double rtol,
                                 int *iters, double *relres, double *solve_ms) {
  if (!ctxs || n_ctx <= 0 || !ctxs[0]) return EMFEM_ERR_ARG;
  emfem_context *c0 = ctxs[0];
  int conv_rc = EMFEM_OK;
  const int rc = guarded(c0, [&]() {
    EMF_REQUIRE(n > 0 && first >= 0 && max_it > 0, kErrArg, "bad batch solve arguments");
    for (int k = 0; k < n_ctx; ++k) {
      emfem_context *c = ctxs[k];
      EMF_REQUIRE(c != nullptr, kErrArg, "null context in batch");
      require_system(c);
      EMF_REQUIRE(c->assembled_fidx >= 0, kErrState, "matrix of batch model " + std::to_string(k) + " not assembled");
      EMF_REQUIRE(first + n <= c->nslots, kErrArg, "slot range out of bounds for batch model " + std::to_string(k));
    }
    // Stages: each solves the models still running until half of their right-hand
    // sides have passed the stopping test; models whose right-hand sides have all
    // passed retire, the others are restacked and continue from their iterates, so
    // every later stage runs on a smaller system (-batch_retire false: one stage
    // in lock-step to completion).
    const bool retire = c0->opt("batch_retire") == "true";
    const int nl = n_ctx * n;
    std::vector<int> its(nl, 0), conv(nl, 0);
    std::vector<double> rr(nl, 0.0), done(nl, -1.0);
    std::vector<int> live(n_ctx);
    for (int k = 0; k < n_ctx; ++k) live[k] = k;
    const double t_begin = now_ms();
    double setup = 0.0, solve = 0.0;
    std::map<std::string, double> phase_ms;
    long inner = 0;
    int stages = 0, fell_back = 0;
    bool warm = false, complete = !retire;
    while (!live.empty()) {
      const int ns = (int)live.size();
      int used = 0, open = 0;
      for (int k : live)
        for (int j = 0; j < n; ++j) {
          used = std::max(used, its[k * n + j]);
          open += !conv[k * n + j];
        }
      const int budget = max_it - used;
      // stop once half of the right-hand sides not yet converged have passed
      const int stop = !complete && ns > kMinStageModels && budget > 0 ? std::max(1, open / 2) : 0;
      StageOut so;
      if (budget > 0) solve_stage(ctxs, live, first, n, budget, rtol, warm, stop, t_begin, so);
      ++stages;
      setup += so.setup_ms;
      solve += so.solve_ms;
      for (auto &kv : so.timing) phase_ms[kv.first] += kv.second;
      inner += so.st.inner_its;
      fell_back |= so.st.fell_back;
      std::vector<int> next;
      for (int i = 0; i < ns; ++i) {
        const int k = live[i];
        bool all = budget > 0;
        for (int j = 0; j < n && budget > 0; ++j) all &= so.st.converged[(size_t)i * n + j] != 0;
        for (int j = 0; j < n && budget > 0; ++j) {
          const size_t q = (size_t)i * n + j;
          its[k * n + j] += so.st.its[q];
          rr[k * n + j] = so.st.relres[q];
          // a right-hand side accepted in an earlier stage keeps its acceptance time
          if (!conv[k * n + j]) done[k * n + j] = so.solve_start_ms + so.st.done_ms[q];
          conv[k * n + j] = so.st.converged[q];
        }
        if (!all && stop > 0) next.push_back(k);
      }
      // nothing retired: the remaining models run to completion in the next stage
      if (stop > 0 && (int)next.size() == ns) complete = true;
      live.swap(next);
      warm = true;
      if (stop == 0) break;
    }
    c0->stats["batch_setup_ms"] = setup;
    for (auto &kv : phase_ms) c0->stats["batch_setup_" + kv.first + "_ms"] = kv.second;
    c0->stats["batch_solve_ms"] = solve;
    c0->stats["batch_wall_ms"] = now_ms() - t_begin;
    c0->stats["batch_stages"] = stages;
    c0->stats["batch_inner_its"] = (double)inner;
    c0->stats["batch_fell_back"] = fell_back;
    int outer = 0;
    for (int q = 0; q < nl; ++q) {
      outer = std::max(outer, its[q]);
      if (iters) iters[q] = its[q];
      if (relres) relres[q] = rr[q];
      if (solve_ms) solve_ms[q] = done[q];
      if (!conv[q]) conv_rc = EMFEM_ERR_NOT_CONVERGED;
    }
    c0->stats["batch_outer_its"] = outer;
    if (conv_rc != EMFEM_OK) c0->err = "the stopping test was not met within max_it for a batch model";
  });
  return rc != EMFEM_OK ? rc : conv_rc;
}

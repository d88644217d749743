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

extern "C" int emfem_solve_batch(emfem_context **ctxs, int n_ctx, int first, int n, int max_it, double rtol,
                                 int *iters, double *relres, double *solve_ms) {
  if (!ctxs || n_ctx <= 0 || !ctxs[0]) return EMFEM_ERR_ARG;
  emfem_context *c0 = ctxs[0];
  int conv_rc = EMFEM_OK;
  const int rc = guarded(c0, [&]() {
    EMF_REQUIRE(n > 0 && first >= 0 && max_it > 0, kErrArg, "bad batch solve arguments");
    int64_t ne = 0, nnz = 0, nv = 0;
    std::vector<int> erow0(n_ctx + 1, 0), nzoff(n_ctx + 1, 0), voff(n_ctx + 1, 0);
    for (int k = 0; k < n_ctx; ++k) {
      emfem_context *c = ctxs[k];
      EMF_REQUIRE(c != nullptr, kErrArg, "null context in batch");
      require_system(c);
      EMF_REQUIRE(c->assembled_fidx >= 0, kErrState, "matrix of batch model " + std::to_string(k) + " not assembled");
      EMF_REQUIRE(first + n <= c->nslots, kErrArg, "slot range out of bounds for batch model " + std::to_string(k));
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
    U.pts.reserve((size_t)nv * 3);
    U.edgelist.reserve((size_t)ne * 2);
    U.vertex_bdr.reserve((size_t)nv);
    for (int k = 0; k < n_ctx; ++k) {
      emfem_context *c = ctxs[k];
      const int e0 = erow0[k], z0 = nzoff[k], v0 = voff[k];
      std::vector<int> hrp(c->T.ne + 1);
      std::vector<double2> hK(c->nnz);
      c->d_rp.download(hrp.data(), hrp.size(), c->stream);
      c->d_ci.download(&B.ci[z0], c->nnz, c->stream);
      c->d_K.download(hK.data(), c->nnz, c->stream);
      EMF_HIP(hipStreamSynchronize(c->stream));
      for (int r = 0; r < c->T.ne; ++r) B.rp[e0 + r + 1] = z0 + hrp[r + 1];
      for (int i = 0; i < c->nnz; ++i) {
        B.v[z0 + i] = hK[i].x + hK[i].y;
        B.ci[z0 + i] += e0;
      }
      EMF_HIP(hipMemcpyAsync(K.get() + z0, c->d_K.get(), sizeof(double2) * c->nnz, hipMemcpyDeviceToDevice, s));
      U.pts.insert(U.pts.end(), c->T.pts.begin(), c->T.pts.end());
      for (int v : c->T.edgelist) U.edgelist.push_back(v + v0);
      U.vertex_bdr.insert(U.vertex_bdr.end(), c->T.vertex_bdr.begin(), c->T.vertex_bdr.end());
      // tets (the pattern of the auxiliary Galerkin products)
      for (int v : c->T.tets) U.tets.push_back(v + v0);
      U.nt += c->T.nt;
    }
    rp.upload(B.rp, s);
    ci.upload(B.ci, s);
    // segments: model k owns rows [erow0[k], erow0[k+1]); reduction chunks of at most
    // kChunkRows rows never straddle two models
    std::vector<int> chunk_row, seg_chunk(n_ctx + 1, 0);
    for (int k = 0; k < n_ctx; ++k) {
      seg_chunk[k] = (int)chunk_row.size();
      const int rows = erow0[k + 1] - erow0[k];
      const int nch = std::max(1, (rows + kChunkRows - 1) / kChunkRows);
      for (int j = 0; j < nch; ++j) chunk_row.push_back(erow0[k] + (int)((int64_t)rows * j / nch));
    }
    seg_chunk[n_ctx] = (int)chunk_row.size();
    chunk_row.push_back((int)ne);
    DevBuf<int> d_row0, d_rseg((size_t)ne), d_chunk_row, d_seg_chunk;
    d_row0.upload(erow0, s);
    d_chunk_row.upload(chunk_row, s);
    d_seg_chunk.upload(seg_chunk, s);
    hipLaunchKernelGGL(k_fill_segments, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, (int)ne, d_row0.get(),
                       n_ctx, d_rseg.get());
    EMF_HIP(hipGetLastError());
    Seg sg;
    sg.nseg = n_ctx;
    sg.nchunks = (int)chunk_row.size() - 1;
    sg.rseg = d_rseg.get();
    sg.chunk_row = d_chunk_row.get();
    sg.seg_chunk = d_seg_chunk.get();
    // auxiliary-space set-up on the union (aggregates never cross models: no
    // couplings between blocks) and the inner solver
    PcOptions o = pc_options(c0);
    AmsHost H;
    build_ams(U, B, o.amg, H);
    auto pc = make_inner_pc(c0->opt("inner_precision") == "fp32", H, (int)ne, (int)nnz, rp.get(), ci.get(), K.get(),
                            o, s);
    pc->set_use_graph(c0->opt("use_graph") == "true");
    // gather the right-hand sides, solve, scatter the solutions
    const int nc = 2 * n;
    DevBuf<double> b((size_t)ne * nc), x((size_t)ne * nc);
    for (int k = 0; k < n_ctx; ++k) {
      emfem_context *c = ctxs[k];
      axpby<double, double>(s, c->T.ne, nc, 1.0, c->d_S.get() + 2 * first, 2 * c->nslots, 0.0,
                            b.get() + (size_t)erow0[k] * nc, nc);
    }
    std::vector<double> tol((size_t)n_ctx * n, rtol);
    OuterOptions oo;
    oo.restart = c0->opti("A_ksp_gmres_restart");
    oo.max_it = max_it;
    oo.fp64_fallback = c0->opt("fp64_fallback") == "true";
    SolveStats st;
    EMF_HIP(hipStreamSynchronize(s));
    const double t1 = now_ms();
    fgmres(s, (int)ne, (int)nnz, rp.get(), ci.get(), K.get(), *pc, n, b.get(), x.get(), tol, oo, st, sg);
    for (int k = 0; k < n_ctx; ++k) {
      emfem_context *c = ctxs[k];
      axpby<double, double>(s, c->T.ne, nc, 1.0, x.get() + (size_t)erow0[k] * nc, nc, 0.0,
                            c->d_X.get() + 2 * first, 2 * c->nslots);
    }
    EMF_HIP(hipStreamSynchronize(s));
    const double t2 = now_ms();
    c0->stats["batch_setup_ms"] = t1 - t0;
    c0->stats["batch_solve_ms"] = t2 - t1;
    c0->stats["batch_inner_its"] = (double)st.inner_its;
    c0->stats["batch_outer_its"] = st.outer_its;
    c0->stats["batch_fell_back"] = st.fell_back;
    for (size_t i = 0; i < st.its.size(); ++i) {
      if (iters) iters[i] = st.its[i];
      if (relres) relres[i] = st.relres[i];
      if (solve_ms) solve_ms[i] = st.done_ms[i];
      if (!st.converged[i]) conv_rc = EMFEM_ERR_NOT_CONVERGED;
    }
    if (conv_rc != EMFEM_OK) c0->err = "the stopping test was not met within max_it for a batch model";
  });
  return rc != EMFEM_OK ? rc : conv_rc;
}

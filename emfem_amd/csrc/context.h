// The EMFEM context shared by the C-ABI translation units.
//
// Reference: EMContext (src/em_ctx.h:15-63).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include <complex>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "amg.h"
#include "common.h"
#include "locate.h"
#include "model.h"
#include "solver.h"
#include "surrogate.h"
#include "topology.h"
#include "fe.h"
#include "kernels.h"
#include "../../include/emfem_amd.h"

using namespace emfem;

// Device copies of the mesh-only matrices of the auxiliary-space Galerkin products
// (for `aux`, held so that the cache cannot outlive it), and the four product value
// blocks.
struct DevGalerkin {
  std::shared_ptr<const AmsAux> aux;
  DevBuf<int> rrp, rci, prp, pci, arp, aci, missing;
  DevBuf<double> rv[4], pv[4], out[4];
  // pinned host staging of the four blocks (page-locked copies run at full link speed)
  double *pinned = nullptr;
  size_t pinned_n = 0;
  DevGalerkin() = default;
  DevGalerkin(const DevGalerkin &) = delete;
  DevGalerkin &operator=(const DevGalerkin &) = delete;
  ~DevGalerkin() {
    if (pinned) (void)hipHostFree(pinned);
  }
  double *staging(size_t n) {
    if (n > pinned_n) {
      if (pinned) EMF_HIP(hipHostFree(pinned));
      pinned = nullptr;
      pinned_n = 0;
      EMF_HIP(hipHostMalloc((void **)&pinned, n * sizeof(double), hipHostMallocDefault));
      pinned_n = n;
    }
    return pinned;
  }
  GalerkinDev view() const {
    GalerkinDev g;
    g.nv0 = aux ? aux->nv0 : 0;
    g.rrp = rrp.get();
    g.rci = rci.get();
    g.prp = prp.get();
    g.pci = pci.get();
    g.arp = arp.get();
    g.aci = aci.get();
    for (int k = 0; k < 4; ++k) {
      g.rv[k] = rv[k].get();
      g.pv[k] = pv[k].get();
    }
    return g;
  }
};

struct emfem_context {
  hipStream_t stream = nullptr;
  bool device = false;  // false: host-only context (topology, options, point location)
  std::string err;
  std::map<std::string, std::string> opts, given;

  Topology T;
  bool have_mesh = false;
  int aniso_form = -1, n_rho = 0;
  std::vector<double> rho, s6;
  std::vector<double> freqs, tx, rx;
  Background bg;
  // observations (read_emd, src/em_io.cc:237-284): otype fidx tidx ridx, obs, obserr, rsp
  std::vector<int> obs4;
  std::vector<std::complex<double>> obs, obserr, rsp;
  // the mesh as read (coarse_mesh, src/em_ctx.h:29) for per-task refinement
  std::vector<double> coarse_pts;
  std::vector<int> coarse_tets, coarse_attr;

  // device mesh
  DevBuf<int> d_tets, d_t2e, d_attr, d_nb, d_non, d_colour, d_v2t;
  DevBuf<double> d_pts, d_s6;
  DevBuf<uint8_t> d_ebdr;

  // linear system
  bool have_system = false;
  int nnz = 0, nslots = 0, assembled_fidx = -1;
  DevBuf<int> d_rp, d_ci;
  // the pattern's 16-bit column form (compress_columns), built at the first
  // preconditioner set-up of the linear system: -1 not yet, 0 does not fit, 1 built
  DevBuf<uint16_t> d_ci16;
  DevBuf<int> d_ci16_base;
  int col16_state = -1;
  DevBuf<double2> d_K;
  DevBuf<double> d_S, d_X;  // [E][2 nslots]
  // solve workspace kept across solves: the compact right-hand side and solution
  // blocks and the outer FGMRES's basis (released with the linear system)
  DevBuf<double> d_sb, d_sx;
  FgmresWork fgw;
  // interleaved staging of the split-layout operator and preconditioner entry points
  DevBuf<double> d_li, d_lo;

  // preconditioner
  std::unique_ptr<AmsHost> ams_host;
  // host copy of the matrix pattern (values B = C + M filled per preconditioner
  // set-up) and the mesh-only auxiliary matrices, both kept while the linear system
  // lives
  HostCsr h_B;
  DevGalerkin gal;
  std::shared_ptr<const AmsAux> ams_aux;
  std::unique_ptr<InnerPc> pc;
  // the preconditioner of an earlier matrix of this linear system (parked by
  // emfem_assemble_matrix): not applied, only refreshed by the next emfem_create_pc
  std::unique_ptr<InnerPc> pc_prev;

  // A set-up started ahead of its frequency (emfem_prepare_pc): the matrix of
  // frequency fidx assembled on stream s2 into K, the auxiliary Galerkin products formed
  // there and downloaded into staging, and the host AMG set-up built from them on the
  // worker thread, while the context's stream runs other work.  The matrix is taken
  // by emfem_assemble_matrix(fidx), the host set-up by the emfem_create_pc after it.
  struct Prepared {
    int fidx = -1;
    bool k_taken = false;
    hipEvent_t e0 = nullptr, e1 = nullptr, done = nullptr;
    DevBuf<double2> K;
    DevBuf<double> bsum, out[4];
    DevBuf<int> missing;
    HostPinned<double> staging;
    HostPinned<int> miss;
    PcOptions o;
    std::unique_ptr<AmsHost> host;
    std::thread worker;
    std::exception_ptr err;
    // the MT right-hand sides of fidx (XY, YX; re and im) from a second worker thread,
    // which needs no GPU; mt_ok once they are complete
    std::thread rhs_worker;
    std::vector<double> mt[2][2];
    bool mt_ok = false;
    double started_ms = 0.0, worker_ms = 0.0;
  } prep;
  hipStream_t s2 = nullptr;
  // the dual right-hand side depends only on the mesh and the receivers: kept until
  // either changes
  std::vector<double> dual_re, dual_im;
  bool dual_ok = false;

  std::map<std::string, double> stats;
  double assemble_kernel_ms = 0.0;

  // learned indicator surrogate: network, feature/target standardization
  // (mean[16], std[16], then target mean, std), and whether the last rank-correlation
  // check rejected it
  std::unique_ptr<SurrogateNet> sur;
  std::vector<float> sur_norm;
  bool sur_rejected = false;

  std::string opt(const std::string &k) const {
    auto it = opts.find(k);
    return it == opts.end() ? std::string() : it->second;
  }
  double optd(const std::string &k) const { return std::strtod(opt(k).c_str(), nullptr); }
  int opti(const std::string &k) const { return (int)std::strtol(opt(k).c_str(), nullptr, 10); }
};
using namespace emfem;

namespace emfem_ctx {

// Waits for a set-up started by emfem_prepare_pc (its stream and worker) and forgets it.
inline void discard_prepared(emfem_context *c) {
  auto &P = c->prep;
  if (P.worker.joinable()) P.worker.join();
  if (P.rhs_worker.joinable()) P.rhs_worker.join();
  P.mt_ok = false;
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  P.fidx = -1;
  P.k_taken = false;
  P.host.reset();
  P.err = nullptr;
}

template <typename F>
inline int guarded(emfem_context *ctx, F &&f) {
  if (!ctx) return EMFEM_ERR_ARG;
  try {
    f();
    return EMFEM_OK;
  } catch (const Error &e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return EMFEM_ERR_USER;
  }
}

inline void require_mesh(const emfem_context *c) { EMF_REQUIRE(c->have_mesh, kErrState, "no mesh: call emfem_set_mesh first"); }
inline void require_system(const emfem_context *c) {
  EMF_REQUIRE(c->have_system, kErrState, "no linear system: call emfem_create_linear_system first");
}
inline void require_slot(const emfem_context *c, int slot) {
  require_system(c);
  EMF_REQUIRE(slot >= 0 && slot < c->nslots, kErrArg, "slot " + std::to_string(slot) + " out of range");
}
inline void require_freq(const emfem_context *c, int fidx) {
  EMF_REQUIRE(fidx >= 0 && fidx < (int)c->freqs.size(), kErrArg, "frequency index " + std::to_string(fidx) + " out of range");
}

inline void require_device(const emfem_context *c) {
  EMF_REQUIRE(c->device, kErrHip, "no HIP device available to this context");
}

inline void upload_mesh(emfem_context *c) {
  if (!c->device) return;
  hipStream_t s = c->stream;
  const Topology &T = c->T;
  c->d_tets.upload(T.tets, s);
  c->d_t2e.upload(T.tet2edge, s);
  c->d_attr.upload(T.attr, s);
  c->d_nb.upload(T.neighbor, s);
  c->d_non.upload(T.non, s);
  c->d_colour.upload(T.color_order, s);
  c->d_pts.upload(T.pts, s);
  c->d_ebdr.upload(T.edge_bdr, s);
  c->d_v2t.upload(T.v2t, s);
  EMF_HIP(hipStreamSynchronize(s));
}

// Batch point location: on the device (locate.hip) when the context has one, else
// the host search of a host-only context.
inline Locate locator(emfem_context *c) {
  if (!c->device) return host_locate(c->T);
  return [c](const std::vector<double> &pts) {
    const int n = (int)(pts.size() / 3);
    std::vector<int> out(n);
    if (n == 0) return out;
    hipStream_t s = c->stream;
    DevBuf<double> q;
    q.upload(pts, s);
    DevBuf<int> nearest((size_t)n), queue((size_t)n * kLocateQueue), res((size_t)n);
    locate_points(s, n, q.get(), c->T.nv, c->d_pts.get(), c->d_tets.get(), c->d_nb.get(), c->d_v2t.get(),
                  nearest.get(), queue.get(), res.get());
    res.download(out.data(), (size_t)n, s);
    EMF_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i)
      EMF_REQUIRE(out[i] != -2, kErrUser,
                  "point location queued " + std::to_string(kLocateQueue) + " tets without an answer for point " +
                      std::to_string(i));
    return out;
  };
}

inline void check_attributes(emfem_context *c) {
  if (!c->have_mesh || c->n_rho == 0) return;
  for (int t = 0; t < c->T.nt; ++t)
    EMF_REQUIRE(c->T.attr[t] >= 0 && c->T.attr[t] < c->n_rho, kErrUser,
                "tet " + std::to_string(t) + " has attribute " + std::to_string(c->T.attr[t]) +
                    " outside the resistivity table");
}

// Copies the (re, im) columns of `slot` between a [E][2 nslots] block and a host pair:
// one contiguous transfer to a device staging buffer, then a strided on-device copy
// (a host<->device 2-D copy with 16-byte rows is row-by-row and very slow).
inline void put_slot(emfem_context *c, DevBuf<double> &blk, int slot, const double *re, const double *im) {
  // mostly-zero right-hand sides (the MT lifting, point sources) go up as their
  // nonzero rows only
  std::vector<int> idx;
  std::vector<double2> v;
  const int ne = c->T.ne;
  for (int e = 0; e < ne && (int64_t)idx.size() * 4 < ne; ++e) {
    const double r = re ? re[e] : 0.0, i = im ? im[e] : 0.0;
    if (r != 0.0 || i != 0.0 || std::signbit(r) || std::signbit(i)) {  // -0.0 kept as written
      idx.push_back(e);
      v.push_back(make_double2(r, i));
    }
  }
  if ((int64_t)idx.size() * 4 < ne) {
    DevBuf<int> didx;
    DevBuf<double2> dv;
    didx.upload(idx, c->stream);
    dv.upload(v, c->stream);
    put_rows(c->stream, ne, (int)idx.size(), didx.get(), dv.get(), blk.get() + 2 * slot, 2 * c->nslots);
    EMF_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  std::vector<double> tmp((size_t)c->T.ne * 2);
  for (int e = 0; e < c->T.ne; ++e) {
    tmp[(size_t)e * 2] = re ? re[e] : 0.0;
    tmp[(size_t)e * 2 + 1] = im ? im[e] : 0.0;
  }
  DevBuf<double> stage;
  stage.upload(tmp, c->stream);
  axpby<double, double>(c->stream, c->T.ne, 2, 1.0, stage.get(), 2, 0.0, blk.get() + 2 * slot, 2 * c->nslots);
  EMF_HIP(hipStreamSynchronize(c->stream));
}

inline void get_slot(const emfem_context *c, const DevBuf<double> &blk, int slot, double *re, double *im) {
  std::vector<double> tmp((size_t)c->T.ne * 2);
  DevBuf<double> stage((size_t)c->T.ne * 2);
  axpby<double, double>(c->stream, c->T.ne, 2, 1.0, blk.get() + 2 * slot, 2 * c->nslots, 0.0, stage.get(), 2);
  stage.download(tmp.data(), tmp.size(), c->stream);
  EMF_HIP(hipStreamSynchronize(c->stream));
  for (int e = 0; e < c->T.ne; ++e) {
    if (re) re[e] = tmp[(size_t)e * 2];
    if (im) im[e] = tmp[(size_t)e * 2 + 1];
  }
}

// Solution of a slot as interleaved (re, im) per edge on the host.
inline std::vector<double> solution_ri(const emfem_context *c, int slot) {
  std::vector<double> out((size_t)c->T.ne * 2);
  DevBuf<double> stage((size_t)c->T.ne * 2);
  axpby<double, double>(c->stream, c->T.ne, 2, 1.0, c->d_X.get() + 2 * slot, 2 * c->nslots, 0.0, stage.get(), 2);
  stage.download(out.data(), out.size(), c->stream);
  EMF_HIP(hipStreamSynchronize(c->stream));
  return out;
}

inline PcOptions pc_options(const emfem_context *c) {
  PcOptions o;
  o.cheb_degree = c->opti("ams_cheb_degree");
  o.cheb_ratio = c->optd("ams_cheb_ratio");
  o.cheb_fourth = c->opt("ams_cheb_kind") == "fourth";
  o.inner_rtol = c->optd("B_ksp_rtol");
  o.inner_max_it = c->opti("B_ksp_max_it");
  o.amg.theta = c->optd("ams_theta");
  o.amg.smooth_levels = c->opti("ams_smooth_levels");
  o.amg.max_coarse = c->opti("ams_max_coarse");
  EMF_REQUIRE(o.amg.max_coarse >= 1 && o.amg.max_coarse <= kDenseMax, kErrArg,
              "ams_max_coarse must lie in [1, " + std::to_string(kDenseMax) + "]");
  o.amg.additive = c->opt("ams_cycle") == "additive";
  o.use_graph = c->opt("use_graph") == "true";
  o.aux_csr = c->opt("ams_transfer") == "csr";
  o.col16 = c->opt("spmm_col16") == "true";
  // spmm_rows: 0 tuned, 1 lanes per row tuned, 2 row blocks, 3 x-blocks, 4 persistent
  // pipelined row blocks, 10 + n fixed n lanes per row
  const std::string sr = c->opt("spmm_rows");
  o.spmm_rows = sr == "lanes" ? 1
                : sr == "rowblock" ? 2
                : sr == "xblock"   ? 3
                : sr == "pipelined" ? 4
                : sr == "auto"     ? 0
                                   : 10 + c->opti("spmm_rows");
  return o;
}

// Learned indicator for the current mesh and solutions (surrogate_ctx.hip).
std::vector<double> surrogate_indicator(emfem_context *c, int fidx, int sxy, int syx);
// Spearman rank correlation of two equally long arrays.
double rank_correlation(const std::vector<double> &a, const std::vector<double> &b);

}  // namespace emfem_ctx

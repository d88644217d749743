// The split-layout operator for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// emfem_matshell_mult_a_split stages the VecNest {re, im} halves through two
// interleaved blocks that the context allocates on first use.  This variant stages
// them in a fixed, file-scope device array instead, so no call allocates or frees.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/emfem_amd.h"
#include "common.h"
#include "context.h"
#include "kernels.h"

using namespace emfem;
using namespace emfem_ctx;

namespace {

// Interleaved staging of the input and the output blocks, kScratchDoubles / 2
// doubles each (32 MiB each).  Each GPU holds its own copy until the process ends.
constexpr size_t kScratchDoubles = size_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per device.
// The array's name is a host-side stand-in, not a device address: only this pointer
// may be handed to kernels.
double *scratch_on_current_device() {
  int dev = 0;
  EMF_HIP(hipGetDevice(&dev));
  EMF_REQUIRE(dev >= 0 && dev < kMaxDevices, kErrHip, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    EMF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EMF_REQUIRE(p != nullptr, kErrHip, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int emfem_matshell_mult_a_split_noalloc(emfem_context *c, int n_rhs, const double *x_re, const double *x_im,
                                        double *y_re, double *y_im) {
  return guarded(c, [&]() {
    require_system(c);
    EMF_REQUIRE(n_rhs > 0 && x_re && x_im && y_re && y_im, kErrArg, "bad operator arguments");
    const int ne = c->T.ne;
    const size_t half = kScratchDoubles / 2;
    const size_t need = (size_t)ne * 2 * (size_t)n_rhs;
    EMF_REQUIRE(need <= half, kErrArg,
                "n_edges * n_rhs = " + std::to_string((size_t)ne * n_rhs) + " exceeds the " +
                    std::to_string(half / 2) + " this entry point can stage without allocating");
    double *in = scratch_on_current_device();
    double *out = in + half;
    hipStream_t s = c->stream;
    split_to_interleaved(s, ne, n_rhs, x_re, x_im, in);
    op_spmm<double>(s, ne, c->nnz, c->d_rp.get(), c->d_ci.get(), c->d_K.get(), n_rhs, in, 2 * n_rhs, out,
                    2 * n_rhs, 1.0, 0.0, nullptr, 0);
    interleaved_to_split(s, ne, n_rhs, out, y_re, y_im);
    EMF_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// Shared host utilities: error type, HIP checks, device buffers, a thread pool loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace emfem {

// Errors travel as exceptions inside the library and become int status codes at
// the C boundary (the reference's SETERRQ/PetscCall convention).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

constexpr int kErrUser = -1;
constexpr int kErrHip = -2;
constexpr int kErrNotConverged = -3;
constexpr int kErrState = -4;
constexpr int kErrArg = -5;

#define EMF_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::emfem::Error(::emfem::kErrHip, std::string("HIP error ") + hipGetErrorString(e_) +   \
                                                 " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define EMF_REQUIRE(cond, code, msg)                                                               \
  do {                                                                                             \
    if (!(cond)) throw ::emfem::Error((code), (msg));                                              \
  } while (0)

// Owning device allocation.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) EMF_HIP(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const T *h, size_t n, hipStream_t s) {
    resize(n);
    if (n) EMF_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <typename A>
  void upload(const std::vector<T, A> &h, hipStream_t s) {
    upload(h.data(), h.size(), s);
  }
  void download(T *h, size_t n, hipStream_t s) const {
    if (n) EMF_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n_) EMF_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
  }
  T *get() { return p_; }
  const T *get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

// Pinned host array that kernels write directly, read on the host after the writing
// stream is synchronised: mapped and coherent (uncached on the device), so the
// kernel's stores reach host memory and not only the device caches.
template <typename T>
class HostPinned {
 public:
  HostPinned() = default;
  ~HostPinned() { release(); }
  HostPinned(const HostPinned &) = delete;
  HostPinned &operator=(const HostPinned &) = delete;
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) EMF_HIP(hipHostMalloc((void **)&p_, n * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *get() { return p_; }
  T &operator[](size_t i) { return p_[i]; }

 private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

inline double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Upper bound on the threads of the host loops started from the calling thread (0 =
// none): a set-up running beside a solve leaves cores to the thread that drives it.
inline int &host_thread_cap() {
  static thread_local int cap = 0;
  return cap;
}

inline int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  int t = (int)std::max(1u, std::min(n, 16u));
  if (host_thread_cap() > 0) t = std::min(t, host_thread_cap());
  return t;
}

// Static-partition parallel loop over [0, n): f(begin, end, thread_index).
template <typename F>
void parallel_ranges(int64_t n, F &&f, int nthreads = 0) {
  if (nthreads <= 0) nthreads = host_threads();
  if (n < 4096 || nthreads == 1) {
    f(int64_t(0), n, 0);
    return;
  }
  nthreads = (int)std::min<int64_t>(nthreads, n / 1024 + 1);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
    th.emplace_back([&, b, e, t]() { f(b, e, t); });
  }
  for (auto &x : th) x.join();
}

}  // namespace emfem

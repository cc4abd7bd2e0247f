// common.h -- shared host/device helpers of the CATHY gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cg {

// Maximum entries of one row of the full CSR pattern.  P1 tets of CATHY's
// extruded meshes give at most ~20 (checked at set-up).
constexpr int kMaxRow = 32;
// Work-group size of row/node kernels (4 wavefronts of 64 lanes).
constexpr int kBlock = 256;
// Upper bound on work-groups of reduction kernels: partial sums are
// written per work-group and summed in a fixed order (deterministic).
constexpr int kMaxRedBlocks = 8192;
// Work-groups of the PCG vector kernels (alpha, dot, beta): a grid-stride
// loop over the rows, one partial sum per group.
constexpr int kVecGroups = 512;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CG_HIP(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      throw ::cg::Error(-2, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CG_REQUIRE(cond, msg)                                              \
  do {                                                                     \
    if (!(cond)) throw ::cg::Error(-1, msg);                               \
  } while (0)

// Pinned host buffer (hipHostMalloc): device-to-host copies into it are
// plain DMA, with no staging through a bounce buffer and no host copy.
template <typename T>
struct Pinned {
  T *p = nullptr;
  size_t n = 0;
  Pinned() = default;
  Pinned(const Pinned &) = delete;
  Pinned &operator=(const Pinned &) = delete;
  void alloc(size_t count) {
    release();
    CG_HIP(hipHostMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1)));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~Pinned() { release(); }
};

// Owning device buffer.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      hipError_t e = hipMalloc(&p, count * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        throw Error(-3, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
      }
    }
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count, hipStream_t s) {
    CG_REQUIRE(count <= n, "upload larger than buffer");
    if (count) CG_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &h, hipStream_t s) {
    if (n < h.size()) alloc(h.size());
    upload(h.data(), h.size(), s);
  }
  void download(T *h, size_t count, hipStream_t s) const {
    CG_REQUIRE(count <= n, "download larger than buffer");
    if (count) CG_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Last error text of the calling thread (cathy_gpu_last_error).
void set_last_error(const std::string &m);

// Runs one C entry: exceptions become the entry's negative status code.
template <typename F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return -3;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
}

inline int grid_for(size_t work, int block = kBlock, int cap = 1 << 30) {
  size_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (size_t)cap) g = cap;
  return (int)g;
}

}  // namespace cg

// Internal helpers shared by the host orchestration and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>

#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpurify.h"

namespace gp {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GP_ERR_OOM
                                                                           : GP_ERR_INTERNAL;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}
#define GP_HIP(x) ::gp::hip_check((x), #x)
#define GP_LAUNCH_CHECK() ::gp::hip_check(hipGetLastError(), "kernel launch")

// Device-memory accounting: every allocation goes through DBuf so the
// optional GPURIFY_MAX_DEVICE_BYTES cap (used by tests to force the OOM
// path from the host) is enforced in one place.
size_t &device_bytes_in_use();
size_t device_bytes_cap();
// Held by every DBuf allocation, free and upload, and by a thread while it
// captures a HIP graph: a batched solve sets up its next group of problems
// on a second thread, whose device allocations and copies must not fall
// inside the solving thread's capture.
std::recursive_mutex &device_mem_mutex();

// Device memory of the library's buffers, cached per device: a released
// block waits in a pending list until the API call that released it has
// ended with the device idle (device_recycle), then serves later requests
// of its size (a block up to a quarter larger, or 1 MiB, than asked is
// reused); nothing is returned to the driver until device_trim (context
// destruction, or a failed hipMalloc, which trims and retries once).  A
// solve that rebuilds buffers of the same sizes call after call (a batch of
// nets, a benchmark step) therefore maps its device memory once, and no
// release inside a call synchronises the device as hipFree does.  All four
// under device_mem_mutex; device_alloc throws GP_ERR_OOM.
void *device_alloc(size_t bytes);
void device_release(void *p);
void device_recycle();
void device_trim();

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
    release();
    if (count == 0) return;
    size_t bytes = count * sizeof(T);
    size_t cap = device_bytes_cap();
    if (cap && device_bytes_in_use() + bytes > cap)
      throw Error(GP_ERR_OOM, "device allocation of " + std::to_string(bytes) +
                                  " bytes exceeds GPURIFY_MAX_DEVICE_BYTES");
    p = static_cast<T *>(device_alloc(bytes));
    n = count;
    device_bytes_in_use() += bytes;
  }
  void release() {
    if (p) {
      std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
      device_release(p);
      device_bytes_in_use() -= n * sizeof(T);
    }
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  // Completes before returning: callers pass temporaries and locals, and an
  // asynchronous copy from pageable memory may read the host buffer after
  // the call returns.
  void upload(const T *h, size_t count, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
    if (count != n) alloc(count);
    if (n) {
      GP_HIP(hipMemcpyAsync(p, h, bytes(), hipMemcpyHostToDevice, s));
      GP_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void zero(hipStream_t s) {
    if (n) GP_HIP(hipMemsetAsync(p, 0, bytes(), s));
  }
};

// Wall time of the host-to-device copies of the caller's panel data
// (vertices, centroids, areas, conductor ids) in the current API call, per
// thread; cleared by the solve entry points, reported as seconds_upload.
double &upload_seconds();

template <typename T>
DBuf<T> to_device(const std::vector<T> &h, hipStream_t s);
template <typename T>
DBuf<T> to_device(const T *h, size_t count, hipStream_t s);

// to_device for panel data: the copy's wall time counts as input upload.
template <typename T>
DBuf<T> upload_panels(const T *h, size_t count, hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  DBuf<T> d = to_device(h, count, s);  // synchronous: upload() waits for the copy
  upload_seconds() += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return d;
}
template <typename T>
DBuf<T> upload_panels(const std::vector<T> &h, hipStream_t s) {
  return upload_panels(h.data(), h.size(), s);
}

template <typename T>
DBuf<T> to_device(const std::vector<T> &h, hipStream_t s) {
  DBuf<T> d;
  d.upload(h, s);
  return d;
}

template <typename T>
DBuf<T> to_device(const T *h, size_t count, hipStream_t s) {
  DBuf<T> d;
  d.upload(h, count, s);
  return d;
}

// Kernel-level event profiler.  When enabled on a context, every launch
// wrapped in a ProfScope records a HIP event pair on the launch stream plus
// the launch's algorithmic flops and bytes; gp_profile_report() sums them per
// kernel name.  Disabled, a ProfScope costs one branch.
struct Prof {
  bool on = false;
  struct Rec {
    const char *name;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    hipEvent_t e;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else {
      hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    return e;
  }
  ~Prof() {
    for (auto &r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
Prof *&active_prof();  // per thread, set by the API entry for its context

struct ProfScope {
  Prof *p;
  hipStream_t s;
  size_t idx = 0;
  ProfScope(const char *name, hipStream_t st, double flops, double bytes) : p(active_prof()), s(st) {
    if (!p || !p->on) {
      p = nullptr;
      return;
    }
    Prof::Rec r{name, p->get(), p->get(), flops, bytes};
    hip_check(hipEventRecord(r.a, s), "hipEventRecord");
    p->recs.push_back(r);
    idx = p->recs.size() - 1;
  }
  ~ProfScope() {
    if (p) (void)hipEventRecord(p->recs[idx].b, s);
  }
};

// Kernel-variant autotuner, one per context.  The first launch of a tunable
// kernel with a given shape key runs each of its variants on the real
// operands (every variant writes the same outputs) and times them with HIP
// events on the launch stream, after one untimed warm-up launch; the fastest
// then serves every later launch with that key (and is launched once more so
// the outputs are its own).  A stream under graph capture launches the
// default variant 0 of an untuned key.  force >= 0 pins every tunable kernel
// to that variant (clamped to its count), for tests.
struct Tuner {
  int force = -1;
  struct Entry {
    std::string kernel;
    int64_t k0, k1, k2;
    int variant;
    std::vector<float> ms;
  };
  std::vector<Entry> table;
};
Tuner *&active_tuner();  // per thread, set by the API entry for its context

template <typename F>
void tuned_launch(const char *kernel, int64_t k0, int64_t k1, int64_t k2, int nvar, hipStream_t s,
                  F &&launch) {
  Tuner *t = active_tuner();
  if (!t || nvar <= 1) {
    launch(0);
    return;
  }
  if (t->force >= 0) {
    launch(t->force < nvar ? t->force : nvar - 1);
    return;
  }
  for (const auto &e : t->table)
    if (e.kernel == kernel && e.k0 == k0 && e.k1 == k1 && e.k2 == k2) {
      launch(e.variant);
      return;
    }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
  if (cs != hipStreamCaptureStatusNone) {
    launch(0);
    return;
  }
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  std::vector<float> ms(nvar, 0.f);
  launch(0);  // warm-up: first touch of the operands
  for (int v = 0; v < nvar; ++v) {
    hip_check(hipEventRecord(a, s), "hipEventRecord");
    launch(v);
    hip_check(hipEventRecord(b, s), "hipEventRecord");
    hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms[v], a, b), "hipEventElapsedTime");
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  int best = 0;
  for (int v = 1; v < nvar; ++v)
    if (ms[v] < ms[best]) best = v;
  if (best != nvar - 1) launch(best);
  t->table.push_back({kernel, k0, k1, k2, best, ms});
}

// Physical constants (crates/pex/src/quasistatic/constants.rs:13-20).
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kPi = 3.14159265358979323846;

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// The work-groups of a launch are dealt to the 8 XCDs of the MI355X round
// robin (hardware block b on XCD b % 8), each XCD with its own L2.  xcd_block
// renumbers them so that each XCD runs runs of `chunk` consecutive logical
// blocks (XCD x takes runs x, x + 8, ...): neighbouring work items, which read
// the same inputs (sibling boxes' source multipoles, one leaf's tile for each
// RHS chunk), run at the same time on one XCD and share its L2, while the
// XCDs still interleave along the whole range.  chunk 0 keeps the hardware
// order.  A bijection of [0, nb) (the last partial round keeps its order).
__device__ inline int xcd_block(int b, int nb, int chunk) {
  if (chunk <= 0) return b;
  const int span = 8 * chunk, full = nb / span * span;
  if (b >= full) return b;
  const int x = b & 7, i = b >> 3;
  return ((i / chunk) * 8 + x) * chunk + i % chunk;
}

}  // namespace gp

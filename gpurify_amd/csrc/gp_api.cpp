// C ABI of libgpurify (include/gpurify.h).  Every entry point catches all
// C++ exceptions and turns them into a status code plus a message.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <unordered_map>
#include <mutex>
#include <string>

#include "../../include/gpurify.h"
#include "gp_bridge.h"
#include "gp_common.h"
#include "gp_fastcap.h"
#include "gp_resistance.h"
#include "gp_m2l_cpqr.h"
#include "gp_parallel.h"
#include "gp_noalloc.h"
#include "gp_solver.h"
#include "gp_tree.h"

namespace gp {

size_t &device_bytes_in_use() {
  static size_t used = 0;
  return used;
}

std::recursive_mutex &device_mem_mutex() {
  static std::recursive_mutex m;
  return m;
}

namespace {
struct DevCache {
  std::multimap<size_t, void *> avail;           // cached blocks by size
  std::vector<void *> pending;                   // released during the current calls
  std::unordered_map<void *, size_t> size_of;    // every block of this device
};
std::map<int, DevCache> &dev_caches() {
  static std::map<int, DevCache> c;
  return c;
}
int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) d = 0;
  return d;
}
void trim_one(DevCache &c) {
  for (auto &kv : c.avail) {
    (void)hipFree(kv.second);
    c.size_of.erase(kv.second);
  }
  c.avail.clear();
}
}  // namespace

void *device_alloc(size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
  DevCache &c = dev_caches()[current_device()];
  auto it = c.avail.lower_bound(bytes);
  if (it != c.avail.end() && it->first <= bytes + std::max(bytes / 4, (size_t)1 << 20)) {
    void *p = it->second;
    c.avail.erase(it);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim_one(c);
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(GP_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes: " +
                                hipGetErrorString(e));
  }
  c.size_of[p] = bytes;
  return p;
}

void device_release(void *p) {
  std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
  dev_caches()[current_device()].pending.push_back(p);
}

void device_recycle() {
  std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
  const int cur = current_device();
  for (auto &kv : dev_caches()) {
    DevCache &c = kv.second;
    if (c.pending.empty()) continue;
    // every kernel that used a pending block was queued before its release:
    // once the device is idle none is left
    (void)hipSetDevice(kv.first);
    if (hipDeviceSynchronize() != hipSuccess) {
      (void)hipGetLastError();
      continue;  // keep them pending rather than reuse memory of a failed device
    }
    for (void *p : c.pending) c.avail.emplace(c.size_of[p], p);
    c.pending.clear();
  }
  (void)hipSetDevice(cur);
}

void device_trim() {
  std::lock_guard<std::recursive_mutex> lk(device_mem_mutex());
  trim_one(dev_caches()[current_device()]);
}

double &upload_seconds() {
  static thread_local double t = 0.0;
  return t;
}

Prof *&active_prof() {
  static thread_local Prof *p = nullptr;
  return p;
}

Tuner *&active_tuner() {
  static thread_local Tuner *t = nullptr;
  return t;
}

size_t device_bytes_cap() {
  const char *e = std::getenv("GPURIFY_MAX_DEVICE_BYTES");
  if (!e || !*e) return 0;
  return (size_t)std::strtoull(e, nullptr, 10);
}

}  // namespace gp

namespace {

thread_local std::string g_thread_error = "";

bool force_no_device() {
  const char *e = std::getenv("GPURIFY_FORCE_NO_DEVICE");
  return e && e[0] == '1';
}

// First device whose arch is gfx950, or -1.
int find_gfx950(int ordinal, std::string *name, int *cus) {
  if (force_no_device()) return -1;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return -1;
  if (ordinal < 0 || ordinal >= count) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) return -1;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return -1;
  if (name) *name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
  if (cus) *cus = prop.multiProcessorCount;
  return ordinal;
}

struct ActiveProf {
  explicit ActiveProf(gp_ctx *c) {
    gp::active_prof() = c ? &c->prof : nullptr;
    gp::active_tuner() = c ? &c->tuner : nullptr;
  }
  ~ActiveProf() {
    gp::active_prof() = nullptr;
    gp::active_tuner() = nullptr;
  }
};

// Recycles the device memory released during the call once it has ended
// (gp::device_recycle), whatever its outcome.
struct RecycleAtEnd {
  ~RecycleAtEnd() {
    try {
      gp::device_recycle();
    } catch (...) {
    }
  }
};

template <typename F>
int guarded(gp_ctx *ctx, F &&f) {
  RecycleAtEnd rec;
  ActiveProf ap(ctx);
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return GP_OK;
  } catch (const gp::Error &e) {
    (ctx ? ctx->last_error : g_thread_error) = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    (ctx ? ctx->last_error : g_thread_error) = "host allocation failed";
    return GP_ERR_OOM;
  } catch (const std::exception &e) {
    (ctx ? ctx->last_error : g_thread_error) = std::string("internal error: ") + e.what();
    return GP_ERR_INTERNAL;
  } catch (...) {
    (ctx ? ctx->last_error : g_thread_error) = "internal error";
    return GP_ERR_INTERNAL;
  }
}

}  // namespace

extern "C" {

const char *gp_version(void) { return "gpurify-mi355x 0.1.0 (gfx950)"; }

int gp_device_ready(void) { return find_gfx950(0, nullptr, nullptr) >= 0 ? 1 : 0; }

int gp_ctx_create(gp_ctx **out, uint32_t device_ordinal) {
  if (!out) {
    g_thread_error = "gp_ctx_create: out is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  *out = nullptr;
  return guarded(nullptr, [&] {
    std::string name;
    int cus = 0;
    int dev = find_gfx950((int)device_ordinal, &name, &cus);
    if (dev < 0) throw gp::Error(GP_ERR_NO_DEVICE, "no usable gfx950 device");
    GP_HIP(hipSetDevice(dev));
    gp_ctx *c = new gp_ctx();
    c->device = dev;
    c->device_name = name;
    c->cu_count = cus;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->setup_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->asm_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      if (c->setup_stream) (void)hipStreamDestroy(c->setup_stream);
      delete c;
      gp::hip_check(e, "hipStreamCreateWithFlags");
    }
    gp::noalloc_init(dev);  // the fixed workspace's address on this device
    *out = c;
  });
}

void gp_ctx_destroy(gp_ctx *ctx) {
  if (!ctx) return;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->blas.reset();
    gp::device_recycle();
    gp::device_trim();  // the cached device memory goes back with the context
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->setup_stream) (void)hipStreamDestroy(ctx->setup_stream);
    if (ctx->asm_stream) (void)hipStreamDestroy(ctx->asm_stream);
  }
  delete ctx;
}

const char *gp_last_error(const gp_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : g_thread_error.c_str();
}

int gp_device_synchronize(gp_ctx *ctx) {
  if (!ctx) {
    g_thread_error = "gp_device_synchronize: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    GP_HIP(hipSetDevice(ctx->device));
    GP_HIP(hipDeviceSynchronize());
  });
}

int gp_device_info(gp_ctx *ctx, char *buf, size_t len) {
  if (!ctx || !buf || len == 0) return GP_ERR_INVALID_INPUT;
  std::snprintf(buf, len, "%s, %d CUs, device %d", ctx->device_name.c_str(), ctx->cu_count,
                ctx->device);
  return GP_OK;
}

void gp_cap_opts_default(gp_cap_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = GP_PREC_FP64;
  o->method = GP_METHOD_AUTO;
  o->fmm_order = 4;      // bridge.rs:403
  o->max_level = 4;      // bridge.rs:403
  o->leaf_target = 64;
  o->restart = 200;      // cap/fmm_solver.rs:164
  o->max_restarts = 50;  // cap/fmm_solver.rs:164
  o->rhs_block = 0;
  o->tol = 1.0e-8;       // bridge.rs:403
  o->inner_tol = 1.0e-2;
  o->max_refine = 30;
  o->m2l_inner = GP_M2L_BF16;
  o->m2l_outer = GP_M2L_FP64;
  o->strict = 0;
  o->dense_limit = 500;  // bridge.rs:30
  o->rhs_first = 0;
  o->rhs_count = 0;
  o->inner_passes = 1;
  o->inner_basis = GP_BASIS_FP32;
  o->m2l_rank = 32;
  o->step_graphs = 1;
  o->m2l_compress = 1;  // bbfmm.rs:132-174
  o->fuse_leaf = 1;
  o->skip_converged = 1;
  o->batch_groups = 0;
  o->near_inner = GP_NEAR_FP32;
  o->inner_order = 0;
  o->ground = GP_GROUND_NONE;
  o->ground_z = 0.0;
  o->m2l_tile_order = 0;
}

int gp_cap_solve(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts, double *C,
                 gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_cap_solve: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !C) throw gp::Error(GP_ERR_INVALID_INPUT, "null panels or C");
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    gp::PanelGeom g;
    gp::panel_geometry(*panels, g);
    gp::cap_solve(ctx, g, o, C, report, nullptr);
  });
}

int gp_cap_solve_dense_noalloc(gp_ctx *ctx, const gp_panels *panels, double *C,
                               gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_cap_solve_dense_noalloc: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !C) throw gp::Error(GP_ERR_INVALID_INPUT, "null panels or C");
    gp::active_prof() = nullptr;  // no profiler events: nothing is created
    GP_HIP(hipSetDevice(ctx->device));
    gp::cap_solve_dense_noalloc(ctx, *panels, C, report);
  });
}

int gp_cap_solve_dielectric(gp_ctx *ctx, const gp_panels *panels, const gp_panel_role *roles,
                            double *C, gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_cap_solve_dielectric: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !roles || !C) throw gp::Error(GP_ERR_INVALID_INPUT, "null panels, roles or C");
    if (panels->n_panels == 0) throw gp::Error(GP_ERR_EMPTY, "empty problem");
    if (panels->n_conductors == 0)
      throw gp::Error(GP_ERR_INVALID_INPUT, "a dielectric problem needs at least one conductor");
    if (report) *report = gp_cap_report{};
    GP_HIP(hipSetDevice(ctx->device));
    gp::PanelGeom g;
    gp::panel_geometry(*panels, g);
    gp::solve_dense_roles(ctx, g, roles, C, report);
  });
}

int gp_cap_solve_dielectric_fmm(gp_ctx *ctx, const gp_panels *panels, const gp_panel_role *roles,
                                const gp_cap_opts *opts, double *C, gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_cap_solve_dielectric_fmm: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !roles || !C) throw gp::Error(GP_ERR_INVALID_INPUT, "null panels, roles or C");
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    if (o.ground != GP_GROUND_NONE)
      throw gp::Error(GP_ERR_INVALID_INPUT, "the dielectric solves take no ground plane");
    gp::PanelGeom g;
    gp::panel_geometry(*panels, g);
    gp::solve_dielectric_fmm(ctx, g, roles, o, C, report);
  });
}

int gp_cap_solve_batched(gp_ctx *ctx, const gp_panels *problems, uint32_t n_problems,
                         const gp_cap_opts *opts, double *C, gp_cap_report *reports) {
  if (!ctx) {
    g_thread_error = "gp_cap_solve_batched: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!problems || !C || n_problems == 0)
      throw gp::Error(GP_ERR_INVALID_INPUT, "null problems or C, or no problems");
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    std::vector<gp::PanelGeom> gs(n_problems);
    std::vector<double *> Cp(n_problems);
    size_t off = 0;
    for (uint32_t p = 0; p < n_problems; ++p) {
      Cp[p] = C + off;
      off += (size_t)problems[p].n_conductors * problems[p].n_conductors;
    }
    // one problem a thread; a bad problem's error (the lowest index's, as a
    // sequential loop reports it) is rethrown
    std::vector<std::exception_ptr> errs(n_problems);
    gp::parallel_for(
        (int)n_problems,
        [&](int p) {
          try {
            gp::panel_geometry(problems[p], gs[p], false);
          } catch (...) {
            errs[p] = std::current_exception();
          }
        },
        1);
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);
    gp::cap_solve_batched(ctx, gs, o, Cp, reports);
  });
}

int gp_cap_jacobian(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts,
                    const gp_param_map *params, double *C, double *J, gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_cap_jacobian: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !C || !params) throw gp::Error(GP_ERR_INVALID_INPUT, "null panels, C or params");
    if (params->n_params > 0 && (!params->param_off || !J))
      throw gp::Error(GP_ERR_INVALID_INPUT, "null param_off or J");
    if (params->n_params > 0 && params->param_off[params->n_params] > 0 &&
        (!params->panel || !params->dverts))
      throw gp::Error(GP_ERR_INVALID_INPUT, "null parameter entries");
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    if (o.ground != GP_GROUND_NONE)
      throw gp::Error(GP_ERR_INVALID_INPUT, "gp_cap_jacobian takes no ground plane");
    gp::PanelGeom g;
    gp::panel_geometry(*panels, g);
    gp::ParamMap pm;
    pm.n = (int)params->n_params;
    pm.off = params->param_off;
    pm.panel = params->panel;
    pm.dv = params->dverts;
    gp::cap_jacobian(ctx, g, o, pm, C, J, report);
  });
}

int gp_fmm_apply(gp_ctx *ctx, const gp_panels *panels, const gp_cap_opts *opts, uint32_t nr,
                 uint32_t precision, const double *X, double *Y) {
  if (!ctx) {
    g_thread_error = "gp_fmm_apply: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    if (!panels || !X || !Y || nr == 0 || nr > 256)
      throw gp::Error(GP_ERR_INVALID_INPUT, "gp_fmm_apply: bad arguments");
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    if (o.ground != GP_GROUND_NONE)
      throw gp::Error(GP_ERR_INVALID_INPUT, "gp_fmm_apply takes no ground plane");
    gp::PanelGeom g;
    gp::panel_geometry(*panels, g);
    gp::fmm_apply_host(ctx, g, o, (int)nr, X, Y, precision == GP_PREC_MIXED ? 1 : 0);
  });
}

int gp_profile_enable(gp_ctx *ctx, int enable) {
  if (!ctx) return GP_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    GP_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->prof.recs) {
      ctx->prof.pool.push_back(r.a);
      ctx->prof.pool.push_back(r.b);
    }
    ctx->prof.recs.clear();
    ctx->prof.on = enable != 0;
  });
}

int gp_tune_force(gp_ctx *ctx, int variant) {
  if (!ctx) return GP_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    ctx->tuner.force = variant < 0 ? -1 : variant;
    ctx->tuner.table.clear();
  });
}

int gp_tune_report(gp_ctx *ctx, char *json, size_t len) {
  if (!ctx || !json || len == 0) return GP_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    std::string out = "[";
    char buf[256];
    for (size_t i = 0; i < ctx->tuner.table.size(); ++i) {
      const auto &e = ctx->tuner.table[i];
      std::snprintf(buf, sizeof(buf), "%s{\"kernel\": \"%s\", \"key\": [%lld, %lld, %lld], \"variant\": %d, \"ms\": [",
                    i ? ", " : "", e.kernel.c_str(), (long long)e.k0, (long long)e.k1,
                    (long long)e.k2, e.variant);
      out += buf;
      for (size_t v = 0; v < e.ms.size(); ++v) {
        std::snprintf(buf, sizeof(buf), "%s%.4f", v ? ", " : "", e.ms[v]);
        out += buf;
      }
      out += "]}";
    }
    out += "]";
    if (out.size() + 1 > len) throw gp::Error(GP_ERR_INVALID_INPUT, "tuning report buffer too small");
    std::memcpy(json, out.c_str(), out.size() + 1);
  });
}

int gp_profile_report(gp_ctx *ctx, char *json, size_t len) {
  if (!ctx || !json || len == 0) return GP_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    GP_HIP(hipStreamSynchronize(ctx->stream));
    struct Agg {
      long n = 0;
      double sec = 0, flops = 0, bytes = 0;
    };
    std::vector<std::pair<std::string, Agg>> agg;
    for (auto &r : ctx->prof.recs) {
      float ms = 0.f;
      GP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto it = std::find_if(agg.begin(), agg.end(), [&](auto &kv) { return kv.first == r.name; });
      if (it == agg.end()) {
        agg.push_back({r.name, Agg()});
        it = agg.end() - 1;
      }
      it->second.n += 1;
      it->second.sec += ms * 1e-3;
      it->second.flops += r.flops;
      it->second.bytes += r.bytes;
      ctx->prof.pool.push_back(r.a);
      ctx->prof.pool.push_back(r.b);
    }
    ctx->prof.recs.clear();
    std::string out = "{";
    char buf[512];
    for (size_t i = 0; i < agg.size(); ++i) {
      std::snprintf(buf, sizeof(buf), "%s\"%s\": {\"launches\": %ld, \"seconds\": %.9g, \"flops\": %.9g, \"bytes\": %.9g}",
                    i ? ", " : "", agg[i].first.c_str(), agg[i].second.n, agg[i].second.sec,
                    agg[i].second.flops, agg[i].second.bytes);
      out += buf;
    }
    out += "}";
    if (out.size() + 1 > len) throw gp::Error(GP_ERR_INVALID_INPUT, "profile report buffer too small");
    std::memcpy(json, out.c_str(), out.size() + 1);
  });
}

int gp_mesh_groups(const gp_volume_group *groups, uint32_t n_groups, double xy_scale,
                   uint32_t refine, uint32_t max_panels, uint32_t *n_panels, double *vertices,
                   uint32_t *conductor, double *panel_eps) {
  return guarded(nullptr, [&] {
    gp::Meshed m;
    gp::mesh_groups(groups, n_groups, xy_scale, (int)refine, max_panels, m);
    size_t n = m.cond.size();
    if (n_panels) *n_panels = (uint32_t)n;
    if (vertices) std::memcpy(vertices, m.verts.data(), n * 12 * sizeof(double));
    if (conductor) std::memcpy(conductor, m.cond.data(), n * sizeof(uint32_t));
    if (panel_eps) std::memcpy(panel_eps, m.eps.data(), n * sizeof(double));
  });
}

int gp_solve_capacitance(gp_ctx *ctx, const gp_volume_group *groups, uint32_t n_groups,
                         double xy_scale, uint32_t refine, uint32_t max_panels,
                         const gp_cap_opts *opts, uint32_t *n_nets, uint32_t *nets_out,
                         double *cap_af_out, double *C_out, gp_cap_report *report) {
  if (!ctx) {
    g_thread_error = "gp_solve_capacitance: ctx is NULL";
    return GP_ERR_INVALID_INPUT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  return guarded(ctx, [&] {
    gp_cap_opts o;
    if (opts)
      o = *opts;
    else
      gp_cap_opts_default(&o);
    GP_HIP(hipSetDevice(ctx->device));
    if (n_groups == 0) {  // bridge.rs:311-313: no volume conductors, no nets
      if (n_nets) *n_nets = 0;
      return;
    }
    gp::Meshed m;
    gp::mesh_groups(groups, n_groups, xy_scale, (int)refine, max_panels, m);
    const int nn = (int)m.nets.size(), nc = nn + 1;
    // A ground plane by images takes the place of the substrate plate, at
    // the plate's height: the conductors alone, the plane as the reference.
    const bool image = o.ground == GP_GROUND_IMAGE;
    if (image) {
      const size_t keep = m.cond.size() - m.n_plate;
      m.cond.resize(keep);
      m.eps.resize(keep);
      m.verts.resize(keep * 12);
      o.ground_z = m.ground_z;
    }
    const int ncs = image ? nn : nc;  // conductors of the solve
    gp_panels p{(uint32_t)m.cond.size(), (uint32_t)ncs, m.verts.data(), m.cond.data()};
    gp::PanelGeom g;
    gp::panel_geometry(p, g);
    // Homogeneity test of bridge.rs:393-396.
    const double first = m.eps[0];
    bool homog = true;
    for (double e : m.eps)
      if (std::fabs(e - first) > std::max(std::fabs(first), 1.0) * 1.0e-12) homog = false;
    std::vector<double> C((size_t)ncs * ncs);
    if (homog) {
      gp::cap_solve(ctx, g, o, C.data(), report, nullptr);
      for (auto &v : C) v *= first;  // bridge.rs:405-407
    } else {
      // Conductor-role dielectric solve (bridge.rs:409-430).  The reference
      // solves it dense and refuses 500 panels or more; GP_METHOD_FMM lifts
      // that limit (the conductor rows are the same potential coefficients,
      // so the FMM operator applies and only the charge weights change).
      gp_cap_opts od = o;
      if (o.method != GP_METHOD_FMM) {
        uint32_t limit = o.dense_limit ? o.dense_limit : 500u;
        if (g.n >= (int)limit)
          throw gp::Error(GP_ERR_PANEL_CAPACITY, "extruded layout needs " + std::to_string(g.n) +
                                                     " panels; limit is " +
                                                     std::to_string(limit - 1));
        od.method = GP_METHOD_DENSE;
      }
      gp::cap_solve(ctx, g, od, C.data(), report, m.eps.data());
    }
    if (image) {
      // the plane's row and column by charge neutrality: it holds minus the
      // conductors' total charge
      std::vector<double> Cf((size_t)nc * nc, 0.0);
      double all = 0.0;
      for (int i = 0; i < nn; ++i)
        for (int j = 0; j < nn; ++j) {
          const double v = C[(size_t)i * nn + j];
          Cf[(size_t)i * nc + j] = v;
          Cf[(size_t)i * nc + nn] -= v;
          Cf[(size_t)nn * nc + j] -= v;
          all += v;
        }
      Cf[(size_t)nn * nc + nn] = all;
      C.swap(Cf);
    }
    if (n_nets) *n_nets = (uint32_t)nn;
    for (int i = 0; i < nn; ++i) {
      double af = C[(size_t)i * nc + i] * 1.0e18;  // bridge.rs:432-441
      if (!std::isfinite(af) || af < 0.0)
        throw gp::Error(GP_ERR_NON_FINITE, "quasi-static solver returned a non-finite capacitance for net " +
                                               std::to_string(m.nets[i]));
      if (nets_out) nets_out[i] = m.nets[i];
      if (cap_af_out) cap_af_out[i] = af;
    }
    if (C_out) std::memcpy(C_out, C.data(), C.size() * sizeof(double));
  });
}


int gp_m2l_operator(uint32_t order, int32_t o0, int32_t o1, int32_t o2, uint32_t compressed,
                    double *G, uint32_t *rank) {
  return guarded(nullptr, [&] {
    if (!G) throw gp::Error(GP_ERR_INVALID_INPUT, "G is NULL");
    if (order < 2 || order > 6) throw gp::Error(GP_ERR_INVALID_INPUT, "order must be in [2, 6]");
    const int a0 = std::abs(o0), a1 = std::abs(o1), a2 = std::abs(o2);
    if (std::max(a0, std::max(a1, a2)) > 3 || std::max(a0, std::max(a1, a2)) < 2)
      throw gp::Error(GP_ERR_INVALID_INPUT, "offset must be a far-field offset in [-3, 3]^3");
    const int n = (int)order, n3 = n * n * n, n3p = (n3 + 31) / 32 * 32;
    std::vector<int> ranks;
    const auto tab = gp::m2l_tables(n, n3p, compressed != 0, &ranks);
    const int oi = (o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3);
    const double *blk = tab->data() + (size_t)oi * n3p * n3p;
    for (int m = 0; m < n3; ++m)
      std::memcpy(G + (size_t)m * n3, blk + (size_t)m * n3p, n3 * sizeof(double));
    if (rank) *rank = (uint32_t)ranks[oi];
  });
}

int gp_m2l_factors(uint32_t order, int32_t o0, int32_t o1, int32_t o2, double *U, double *Vt,
                   uint32_t *rank) {
  return guarded(nullptr, [&] {
    if (!U || !Vt) throw gp::Error(GP_ERR_INVALID_INPUT, "U or Vt is NULL");
    if (order < 2 || order > 6) throw gp::Error(GP_ERR_INVALID_INPUT, "order must be in [2, 6]");
    const int a0 = std::abs(o0), a1 = std::abs(o1), a2 = std::abs(o2);
    if (std::max(a0, std::max(a1, a2)) > 3 || std::max(a0, std::max(a1, a2)) < 2)
      throw gp::Error(GP_ERR_INVALID_INPUT, "offset must be a far-field offset in [-3, 3]^3");
    const int n = (int)order, n3 = n * n * n, n3p = (n3 + 31) / 32 * 32;
    const auto f = gp::m2l_factors(n, n3p);
    const int oi = (o0 + 3) * 49 + (o1 + 3) * 7 + (o2 + 3), k = f->rank[oi], kp = f->kp;
    std::fill(U, U + (size_t)n3 * n3, 0.0);
    std::fill(Vt, Vt + (size_t)n3 * n3, 0.0);
    for (int s = 0; s < k; ++s) {
      for (int l = 0; l < n3; ++l) Vt[(size_t)s * n3 + l] = f->Vt[((size_t)oi * kp + s) * n3p + l];
      for (int m = 0; m < n3; ++m)
        U[(size_t)m * n3 + s] = f->Up[(((size_t)oi * n3p + m) * 4 + s % 4) * (kp / 4) + s / 4];
    }
    if (rank) *rank = (uint32_t)k;
  });
}

int gp_qui_parse(const char *text, size_t text_len, double scale, uint32_t *n_panels,
                 uint32_t *n_conductors, size_t *names_len, double *vertices,
                 uint32_t *conductor, uint8_t *n_vertices, char *names) {
  return guarded(nullptr, [&] {
    if (!text && text_len) throw gp::Error(GP_ERR_INVALID_INPUT, "gp_qui_parse: text is NULL");
    gp::QuiGeometry geo;
    gp::parse_qui(text, text_len, scale, geo);
    const size_t n = geo.cond.size();
    size_t nl = 0;
    for (const auto &s : geo.names) nl += s.size() + 1;
    if (n_panels) *n_panels = (uint32_t)n;
    if (n_conductors) *n_conductors = (uint32_t)geo.names.size();
    if (names_len) *names_len = nl;
    if (vertices && n) std::memcpy(vertices, geo.verts.data(), n * 12 * sizeof(double));
    if (conductor && n) std::memcpy(conductor, geo.cond.data(), n * sizeof(uint32_t));
    if (n_vertices && n) std::memcpy(n_vertices, geo.n_vertices.data(), n);
    if (names)
      for (const auto &s : geo.names) {
        std::memcpy(names, s.c_str(), s.size() + 1);
        names += s.size() + 1;
      }
  });
}

int gp_solve_resistance(const gp_res_polygon *polys, uint32_t n_polys, double dbu_nm,
                        uint32_t *n_nets, uint32_t *nets_out, double *r_ohm_out) {
  return guarded(nullptr, [&] {
    std::map<uint32_t, double> r;
    gp::solve_resistance(polys, n_polys, dbu_nm, r);
    if (n_nets) *n_nets = (uint32_t)r.size();
    size_t i = 0;
    for (const auto &e : r) {
      if (nets_out) nets_out[i] = e.first;
      if (r_ohm_out) r_ohm_out[i] = e.second;
      ++i;
    }
  });
}

}  // extern "C"

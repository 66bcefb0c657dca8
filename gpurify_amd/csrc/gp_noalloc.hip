// The dense capacitance solve in a fixed device workspace: the variant of
// gp_cap_solve's dense path (collocation assembly, one-work-group LU,
// charge accumulation) that allocates no device memory, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// The workspace is the file-scope device array `scratch` below; each device
// holds its own copy for the life of the process.  Its address on a device
// comes from hipGetSymbolAddress, taken once per device when the first
// context on that device is created (noalloc_init), and only that address is
// ever handed to kernels and copies.  Host staging lives in one static
// buffer; a process-wide lock serializes the calls, so calls from several
// threads of one process never share the workspace at the same time (calls
// from separate processes use separate copies).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>

#include "../../include/gpurify.h"
#include "gp_common.h"
#include "gp_kernels.h"
#include "gp_noalloc.h"
#include "gp_solver.h"
#include "gp_tree.h"

namespace gp {
namespace {

constexpr int kMaxPanels = GP_NOALLOC_MAX_PANELS;
constexpr int kMaxCond = GP_NOALLOC_MAX_CONDUCTORS;
constexpr size_t a256(size_t b) { return (b + 255) / 256 * 256; }
// Workspace regions (bytes from the start, 256-byte aligned).
constexpr size_t kP = 0;                                                    // n x n FP64
constexpr size_t kVerts = kP + a256(8ull * kMaxPanels * kMaxPanels);        // n x 12
constexpr size_t kCen = kVerts + a256(8ull * 12 * kMaxPanels);              // n x 3
constexpr size_t kW = kCen + a256(8ull * 3 * kMaxPanels);                   // n
constexpr size_t kB = kW + a256(8ull * kMaxPanels);                         // n x nc
constexpr size_t kWork = kB + a256(8ull * kMaxPanels * kMaxCond);           // n x nc
constexpr size_t kC = kWork + a256(8ull * kMaxPanels * kMaxCond);           // nc x nc
constexpr size_t kCond = kC + a256(8ull * kMaxCond * kMaxCond);             // n u32
constexpr size_t kCoff = kCond + a256(4ull * kMaxPanels);                   // nc + 1 int
constexpr size_t kCidx = kCoff + a256(4ull * (kMaxCond + 1));               // n int
constexpr size_t kPiv = kCidx + a256(4ull * kMaxPanels);                    // n int
constexpr size_t kInfo = kPiv + a256(4ull * kMaxPanels);                    // 1 int
constexpr size_t kBytes = kInfo + 256;
static_assert(kBytes == GP_NOALLOC_SCRATCH_BYTES, "workspace size mismatch");

constexpr int kMaxDevices = 64;

}  // namespace
}  // namespace gp

static __device__ __align__(256) unsigned char scratch[gp::kBytes];

namespace gp {
namespace {

std::mutex g_dev_mu;
bool g_dev_done[kMaxDevices] = {};
hipError_t g_dev_err[kMaxDevices] = {};
unsigned char *g_dev_ptr[kMaxDevices] = {};

// Host staging of one call (static: the entry point allocates nothing).
struct Stage {
  double verts[12 * kMaxPanels], cen[3 * kMaxPanels], w[kMaxPanels];
  double cx[kMaxPanels], cy[kMaxPanels], cz[kMaxPanels];
  uint32_t cond[kMaxPanels];
  int coff[kMaxCond + 1], cidx[kMaxPanels], fill[kMaxCond];
  double C[kMaxCond * kMaxCond];
  int info;
};
std::mutex g_stage_mu;
Stage g_stage;

}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_dev_done[device]) return;
  // the caller made `device` current
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_dev_err[device] = e;
  g_dev_ptr[device] = e == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
  g_dev_done[device] = true;
}

void cap_solve_dense_noalloc(gp_ctx *ctx, const gp_panels &in, double *C, gp_cap_report *rep) {
  const double t0 = [] {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
  }();
  if (in.n_panels == 0) throw Error(GP_ERR_EMPTY, "no panels in geometry");
  if (!in.vertices || !in.conductor) throw Error(GP_ERR_INVALID_INPUT, "null panel arrays");
  if (in.n_conductors == 0) throw Error(GP_ERR_INVALID_INPUT, "n_conductors must be > 0");
  if (in.n_panels > (uint32_t)kMaxPanels || in.n_conductors > (uint32_t)kMaxCond)
    throw Error(GP_ERR_INVALID_INPUT,
                "gp_cap_solve_dense_noalloc takes at most " + std::to_string(kMaxPanels) +
                    " panels and " + std::to_string(kMaxCond) + " conductors (got " +
                    std::to_string(in.n_panels) + " and " + std::to_string(in.n_conductors) + ")");
  const int dev = ctx->device;
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (dev < 0 || dev >= kMaxDevices || !g_dev_done[dev])
      throw Error(GP_ERR_INTERNAL, "fixed workspace of device " + std::to_string(dev) +
                                       " was not set up");
    if (g_dev_err[dev] != hipSuccess)
      hip_check(g_dev_err[dev], "hipGetSymbolAddress(scratch)");
    base = g_dev_ptr[dev];
  }
  std::lock_guard<std::mutex> lk(g_stage_mu);
  Stage &h = g_stage;
  const int n = (int)in.n_panels, nc = (int)in.n_conductors;
  // Panel checks and geometry exactly as gp_cap_solve (panel.rs:29-67).
  for (int i = 0; i < n; ++i) {
    check_panel(in, i);
    double c[3];
    if (!panel_props(in.vertices + (size_t)i * 12, c, h.w[i]))
      throw Error(GP_ERR_INVALID_INPUT, "panel " + std::to_string(i) + " has zero area");
    h.cx[i] = c[0];
    h.cy[i] = c[1];
    h.cz[i] = c[2];
  }
  double c[3], L0;
  root_cube(n, h.cx, h.cy, h.cz, c, L0);
  // Unit frame: lengths / L0 about the root centre (the dense path's scaling).
  const double inv = 1.0 / L0;
  for (int p = 0; p < n; ++p) {
    for (int k = 0; k < 12; ++k)
      h.verts[p * 12 + k] = (in.vertices[(size_t)p * 12 + k] - c[k % 3]) * inv;
    h.cen[3 * p] = (h.cx[p] - c[0]) * inv;
    h.cen[3 * p + 1] = (h.cy[p] - c[1]) * inv;
    h.cen[3 * p + 2] = (h.cz[p] - c[2]) * inv;
    h.w[p] = h.w[p] * inv * inv;
    h.cond[p] = in.conductor[p];
  }
  // Panels of each conductor in ascending order (accumulate_c, cap/solver.rs:67-71).
  for (int j = 0; j <= nc; ++j) h.coff[j] = 0;
  for (int i = 0; i < n; ++i) h.coff[h.cond[i] + 1]++;
  for (int j = 0; j < nc; ++j) h.coff[j + 1] += h.coff[j];
  for (int j = 0; j < nc; ++j) h.fill[j] = h.coff[j];
  for (int i = 0; i < n; ++i) h.cidx[h.fill[h.cond[i]]++] = i;

  hipStream_t s = ctx->stream;
  double *P = reinterpret_cast<double *>(base + kP);
  double *verts = reinterpret_cast<double *>(base + kVerts);
  double *cen = reinterpret_cast<double *>(base + kCen);
  double *w = reinterpret_cast<double *>(base + kW);
  double *B = reinterpret_cast<double *>(base + kB);
  double *work = reinterpret_cast<double *>(base + kWork);
  double *Cd = reinterpret_cast<double *>(base + kC);
  uint32_t *cond = reinterpret_cast<uint32_t *>(base + kCond);
  int *coff = reinterpret_cast<int *>(base + kCoff);
  int *cidx = reinterpret_cast<int *>(base + kCidx);
  int *piv = reinterpret_cast<int *>(base + kPiv);
  int *info = reinterpret_cast<int *>(base + kInfo);
  GP_HIP(hipMemcpyAsync(verts, h.verts, sizeof(double) * 12 * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(cen, h.cen, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(w, h.w, sizeof(double) * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(cond, h.cond, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(coff, h.coff, sizeof(int) * (nc + 1), hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(cidx, h.cidx, sizeof(int) * n, hipMemcpyHostToDevice, s));
  // Every region the kernels read is written first in this call: the
  // assembly writes all of P, the RHS kernel all of B, the LU all of work
  // and info (0, or the singular column), the accumulation all of C.
  launch_dense_assemble(n, verts, cen, P, s);
  kr_unit_rhs(cond, 0, n, nc, nc, B, s);
  launch_lu_solve(n, nc, P, piv, B, work, info, s);
  launch_cap_accumulate(nc, nc, nc, coff, cidx, w, B, Cd, s);
  GP_HIP(hipMemcpyAsync(&h.info, info, sizeof(int), hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(h.C, Cd, sizeof(double) * nc * nc, hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  if (h.info != 0) throw Error(GP_ERR_SINGULAR, "potential-coefficient matrix is singular");
  const double scale = 4.0 * kPi * kEps0 * L0;
  for (int e = 0; e < nc * nc; ++e) {
    C[e] = h.C[e] * scale;
    if (!std::isfinite(C[e])) throw Error(GP_ERR_NON_FINITE, "solver returned a non-finite capacitance");
  }
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->n_panels = (uint32_t)n;
    rep->n_conductors = (uint32_t)nc;
    rep->method_used = GP_METHOD_DENSE;
    rep->precision_used = GP_PREC_FP64;
    rep->converged = 1;
    rep->final_rel_residual = 0.0;
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    rep->seconds_total = rep->seconds_solve = ts.tv_sec + 1e-9 * ts.tv_nsec - t0;
  }
}

}  // namespace gp

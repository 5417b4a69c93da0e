// C-ABI entry points: context, step and evolve orchestration, standalone operators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/jathenax_amd.h"
#include "jx_aux.hpp"
#include "jx_debug.hpp"
#include "jx_batch_host.hpp"
#include "jx_stage.hpp"

jx::BatchBase* jx_make_batch_f32(int device, int max_inst,
                                 long long pool_bytes, std::string* err);
jx::BatchBase* jx_make_batch_f64(int device, int max_inst,
                                 long long pool_bytes, std::string* err);

struct jx_batch {
  jx::BatchBase* impl;
  std::string err;
};

// Static scratch of jx_isothermal_step_noalloc: device memory that needs no allocation at
// call time.  Each GPU holds its own copy until the process ends.  Layout: the kick's
// per-block partial sums (the first kNoAllocPartialBytes), then the stage registers B, C
// and, for RK tables with a delta, S, each rounded up to 256 bytes.  Host code never uses
// the array's name as an address: jx_ctx_create takes each device's address once with
// hipGetSymbolAddress (scratch_setup), and only that address reaches kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)96 << 20;
constexpr size_t kNoAllocPartialBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

using jx::Ctrl;
using jx::StageArgs;

// Stage launchers, one per (precision, reconstruction) translation unit.
#define JX_DECL(T, NAME)                                                                    \
  hipError_t NAME(int rs, int nv, bool visc, const jx::StageArgs<T>& a, dim3 grid, int bx, \
                  hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);                         \
  int NAME##_occupancy(int rs, int nv, bool visc, int bx, bool exact);
JX_DECL(double, jx_stage_f64_dc)
JX_DECL(double, jx_stage_f64_plm)
JX_DECL(double, jx_stage_f64_ppm4)
JX_DECL(double, jx_stage_f64_ppmx)
JX_DECL(double, jx_stage_f64_wenoz)
JX_DECL(float, jx_stage_f32_dc)
JX_DECL(float, jx_stage_f32_plm)
JX_DECL(float, jx_stage_f32_ppm4)
JX_DECL(float, jx_stage_f32_ppmx)
JX_DECL(float, jx_stage_f32_wenoz)
#undef JX_DECL

struct jx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Ctrl* ctrl = nullptr;       // device control block
  Ctrl* host_ctrl = nullptr;  // pinned mirror for readback
  void* buf[3] = {nullptr, nullptr, nullptr};   // ping-pong registers B, C and saved S
  size_t buf_bytes = 0;
  void* raw = nullptr;        // rendered raw force (evolve with forcing)
  size_t raw_bytes = 0;
  void* groups = nullptr;     // FP32 render: mode groups of the forcing basis
  size_t groups_bytes = 0;
  void* force_pq = nullptr;   // FP32 evaluated forcing: per-row group sums (ny, 4, 4)
  size_t force_pq_bytes = 0;
  bool fused_forcing = true;  // FP32 forced Smagorinsky runs evaluate the force in place
  double* partial = nullptr;  // per-block partial sums
  size_t partial_count = 0;
  double* sums = nullptr;     // small device result vector
  double* host_sums = nullptr;
  int block_x = 0, rows = 0;
  int cus = 0;                // compute units of the device
  std::string err;
  // Optional HIP-event timing of the fused stage launches.
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t events_used = 0;
};

namespace {

// Per-device address of `scratch` (nullptr until set up) and the setup's status.
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchDevices] = {};
hipError_t g_scratch_status[kScratchDevices] = {};

// Once per device, with that device current: the device address of `scratch`.
void scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    g_scratch_addr[device] = p;
  } else {
    g_scratch_status[device] = (e == hipSuccess) ? hipErrorInvalidSymbol : e;
    (void)hipGetLastError();   // the failure is reported by the entry point, not later calls
  }
}

void* scratch_addr(int device, hipError_t* status) {
  if (device < 0 || device >= kScratchDevices) {
    *status = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  *status = g_scratch_status[device];
  return g_scratch_addr[device];
}

int fail(jx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(jx_ctx* c, hipError_t e, const char* where) {
  return fail(c, JX_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define JX_HIP(c, call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail(c, e_, #call); \
  } while (0)

int ensure(jx_ctx* c, void** p, size_t* have, size_t need) {
  if (*have >= need) return JX_OK;
  if (*p) JX_HIP(c, hipFree(*p));
  *p = nullptr;
  *have = 0;
  JX_HIP(c, hipMalloc(p, need));
  *have = need;
  return JX_OK;
}

int ensure_partial(jx_ctx* c, size_t count) {
  size_t have = c->partial_count * sizeof(double);
  void* p = c->partial;
  int rc = ensure(c, &p, &have, count * sizeof(double));
  c->partial = static_cast<double*>(p);
  c->partial_count = have / sizeof(double);
  return rc;
}

int ensure_buffers(jx_ctx* c, size_t bytes) {
  if (c->buf_bytes >= bytes) return JX_OK;
  for (auto& b : c->buf) {
    if (b) JX_HIP(c, hipFree(b));
    b = nullptr;
  }
  c->buf_bytes = 0;
  for (auto& b : c->buf) JX_HIP(c, hipMalloc(&b, bytes));
  c->buf_bytes = bytes;
  return JX_OK;
}

dim3 render_grid(int nx, int ny) {
  const int rb = jx::render_rows_per_block(nx, ny);
  return dim3((unsigned)((nx + jx::kAuxBlock - 1) / jx::kAuxBlock), (unsigned)((ny + rb - 1) / rb));
}

// FP32 renders sum modes by groups that share their x basis rows (render_grouped_kernel);
// the groups are found from the basis tables on the context stream before the first render.
int prepare_groups(jx_ctx* c, const float* cx, const float* sx, int nx, int n_modes) {
  int rc = ensure(c, &c->groups, &c->groups_bytes, jx::mode_group_ints(n_modes) * sizeof(int));
  if (rc) return rc;
  int* grp = static_cast<int*>(c->groups);
  if (n_modes > 1)
    hipLaunchKernelGGL(jx::mode_relation_kernel<float>, dim3(n_modes, n_modes),
                       dim3(jx::kAuxBlock), 0, c->stream, cx, sx, nx, n_modes, grp);
  hipLaunchKernelGGL(jx::mode_groups_kernel, dim3(1), dim3(1), 0, c->stream, n_modes, grp);
  return JX_OK;
}

template <typename T> int prepare_render(jx_ctx* c, const jx::RenderArgs<T>& ra) {
  if constexpr (std::is_same<T, float>::value)
    return prepare_groups(c, ra.cx, ra.sx, ra.nx, ra.n_modes);
  return JX_OK;
}

size_t force_pq_bytes(int ny) { return (size_t)ny * jx::kMaxForceGroups * 4 * sizeof(float); }

// `eval`: the run evaluates the force where it is used (FP32): only the per-row group sums
// are computed.
template <typename T> void launch_render(jx_ctx* c, const jx::RenderArgs<T>& ra,
                                         bool eval = false) {
  if constexpr (std::is_same<T, float>::value) {
    if (eval) {
      const unsigned nt = (unsigned)ra.ny * jx::kMaxForceGroups;
      hipLaunchKernelGGL(jx::force_rows_kernel, dim3((nt + jx::kAuxBlock - 1) / jx::kAuxBlock),
                         dim3(jx::kAuxBlock), 0, c->stream, ra,
                         static_cast<const int*>(c->groups), static_cast<float4*>(c->force_pq));
      return;
    }
    hipLaunchKernelGGL(jx::render_grouped_kernel, render_grid(ra.nx, ra.ny), dim3(jx::kAuxBlock),
                       0, c->stream, ra, static_cast<const int*>(c->groups));
  } else {
    hipLaunchKernelGGL(jx::render_force_kernel<T>, render_grid(ra.nx, ra.ny),
                       dim3(jx::kAuxBlock), 0, c->stream, ra);
  }
}

dim3 aux_grid(size_t n) {
  size_t b = (n + jx::kAuxBlock - 1) / jx::kAuxBlock;
  b = std::max<size_t>(1, std::min<size_t>(b, 4096));
  return dim3((unsigned)b);
}

using jx::RkRow;

const RkRow* builtin_rk_rows(int integrator) {
  return (integrator >= 2 && integrator <= 4) ? jx::kBuiltinRk[integrator - 2] : nullptr;
}

// The rows a run applies: the caller's table (rk_rows, delta NaN = None) when given,
// else the built-in table of `integrator` stages.  Empty on a bad table.
struct RkTable {
  RkRow row[JX_MAX_RK_STAGES];
  int n = 0;
};

RkTable rk_table(const jx_run_params* p) {
  RkTable t;
  if (!p->rk_rows) {
    const RkRow* b = builtin_rk_rows(p->integrator);
    if (!b) return t;
    for (int k = 0; k < p->integrator; ++k) t.row[k] = b[k];
    t.n = p->integrator;
    return t;
  }
  if (p->integrator < 1 || p->integrator > JX_MAX_RK_STAGES) return t;
  for (int k = 0; k < p->integrator; ++k) {
    const double* r = p->rk_rows + 4 * k;
    if (!std::isfinite(r[0]) || !std::isfinite(r[1]) || !std::isfinite(r[2]) ||
        std::isinf(r[3]))
      return RkTable{};
    const bool has_delta = !std::isnan(r[3]);
    t.row[k] = RkRow{r[0], r[1], r[2], has_delta ? r[3] : 0.0, has_delta ? 1 : 0};
  }
  t.n = p->integrator;
  return t;
}

// The first row reads no saved register when it is (0, 1, beta, None): the fused
// first-stage instantiation (ST1) then applies.
bool plain_first_row(const RkRow& r) {
  return r.g0 == 0.0 && r.g1 == 1.0 && !r.has_delta;
}

int check_params(jx_ctx* c, const jx_run_params* p) {
  if (!p) return fail(c, JX_EINVAL, "params must not be NULL");
  if (p->dtype != 0 && p->dtype != 1) return fail(c, JX_EINVAL, "dtype must be 0 (f64) or 1 (f32)");
  if (p->nx < 1 || p->ny < 1) return fail(c, JX_EINVAL, "grid must be non-empty");
  if (p->nvar < 3 || p->nvar > 5)
    return fail(c, JX_EINVAL, "nvar must be 3 (hydro) plus at most two passive scalars");
  if (p->reconstruct < 0 || p->reconstruct > 4)
    return fail(c, JX_EKEY, "unknown reconstruction id");
  if (p->rsolver < 0 || p->rsolver > 2) return fail(c, JX_EKEY, "unknown Riemann solver id");
  if (rk_table(p).n == 0)
    return p->rk_rows ? fail(c, JX_EINVAL, "RK table rows must be finite (delta NaN for None), "
                                           "1 to 8 stages")
                      : fail(c, JX_EKEY, "unknown integrator");
  const double nonneg[3] = {p->viscosity, p->smagorinsky_cs, p->drag_rate};
  const char* names[3] = {"viscosity", "smagorinsky_cs", "drag_rate"};
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(nonneg[k]) || nonneg[k] < 0.0)
      return fail(c, JX_EINVAL, std::string(names[k]) + " must be finite and nonnegative");
  return JX_OK;
}

template <typename T> struct StageLaunch;
template <> struct StageLaunch<double> {
  static int occupancy(int recon, int rs, int nv, bool visc, int bx, bool exact) {
    switch (recon) {
      case jx::RECON_DC: return jx_stage_f64_dc_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PLM: return jx_stage_f64_plm_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PPM4: return jx_stage_f64_ppm4_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PPMX: return jx_stage_f64_ppmx_occupancy(rs, nv, visc, bx, exact);
      default: return jx_stage_f64_wenoz_occupancy(rs, nv, visc, bx, exact);
    }
  }
  static hipError_t go(int recon, int rs, int nv, bool visc, const StageArgs<double>& a,
                       dim3 g, int bx, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    switch (recon) {
      case jx::RECON_DC: return jx_stage_f64_dc(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PLM: return jx_stage_f64_plm(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PPM4: return jx_stage_f64_ppm4(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PPMX: return jx_stage_f64_ppmx(rs, nv, visc, a, g, bx, st, e0, e1);
      default: return jx_stage_f64_wenoz(rs, nv, visc, a, g, bx, st, e0, e1);
    }
  }
};
template <> struct StageLaunch<float> {
  static int occupancy(int recon, int rs, int nv, bool visc, int bx, bool exact) {
    switch (recon) {
      case jx::RECON_DC: return jx_stage_f32_dc_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PLM: return jx_stage_f32_plm_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PPM4: return jx_stage_f32_ppm4_occupancy(rs, nv, visc, bx, exact);
      case jx::RECON_PPMX: return jx_stage_f32_ppmx_occupancy(rs, nv, visc, bx, exact);
      default: return jx_stage_f32_wenoz_occupancy(rs, nv, visc, bx, exact);
    }
  }
  static hipError_t go(int recon, int rs, int nv, bool visc, const StageArgs<float>& a,
                       dim3 g, int bx, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    switch (recon) {
      case jx::RECON_DC: return jx_stage_f32_dc(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PLM: return jx_stage_f32_plm(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PPM4: return jx_stage_f32_ppm4(rs, nv, visc, a, g, bx, st, e0, e1);
      case jx::RECON_PPMX: return jx_stage_f32_ppmx(rs, nv, visc, a, g, bx, st, e0, e1);
      default: return jx_stage_f32_wenoz(rs, nv, visc, a, g, bx, st, e0, e1);
    }
  }
};

bool is_pow2(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return false;
  int e = 0;
  return std::frexp(x, &e) == 0.5 && x >= 1e-300 && x <= 1e300;
}

// Strip width: the whole row when it fits one block, else 128 threads (two waves per
// barrier).  Rows per block: given, or sized so that the grid is one full wave of resident
// blocks (no tail wave), with at least 12 rows: on grids too small to fill the GPU, more
// blocks beat the cost of each block's 6-row warm-up (1024^2: 12 rows run twice as fast
// as 32).
template <typename T>
void pick_tiles(const jx_ctx* c, const jx_run_params* p, bool visc, bool exact, int* bx,
                int* rows) {
  int b = c->block_x;
  if (b <= 0) b = std::min(128, 64 * ((p->nx + 6 + 63) / 64));
  int r = c->rows;
  if (r <= 0) {
    const int strips = (p->nx + (b - 6) - 1) / (b - 6);
    const int per_cu = StageLaunch<T>::occupancy(p->reconstruct, p->rsolver, p->nvar, visc, b,
                                                 exact);
    const long long slots = (long long)std::max(per_cu, 1) * std::max(c->cus, 1);
    const long long per_strip = std::max<long long>(1, slots / strips);
    r = (int)std::max<long long>(12, (p->ny + per_strip - 1) / per_strip);
  }
  *bx = b;
  *rows = std::max(1, std::min(r, p->ny));
}

// Everything the stage launches of one step need.
template <typename T> struct StepPlan {
  const jx_run_params* p;
  bool visc;
  dim3 grid;
  int bx, rows;
  StageArgs<T> base;
};

// Learned-stress hook of a step: before each stage the stage input's primitives are
// written to W and fn fills tau (ny, nx, 3) on the context stream.
template <typename T> struct SgsHook {
  T* W;
  T* tau;
  jx_stress_fn fn;
  void* user;
};

// `sgs`: the stages add a learned stress, which only the viscous kernel variants read.
template <typename T> StepPlan<T> make_plan(jx_ctx* c, const jx_run_params* p,
                                            bool sgs = false) {
  StepPlan<T> s;
  s.p = p;
  const bool visc_terms = p->viscosity > 0.0 || p->smagorinsky_cs > 0.0;
  s.visc = visc_terms || sgs;
  const bool exact = is_pow2(p->dx) && is_pow2(p->dy) && is_pow2(p->cs);
  pick_tiles<T>(c, p, s.visc, exact, &s.bx, &s.rows);
  s.grid = dim3((unsigned)((p->nx + (s.bx - 6) - 1) / (s.bx - 6)),
                (unsigned)((p->ny + s.rows - 1) / s.rows));
  StageArgs<T> a{};
  a.ctrl = c->ctrl;
  a.nx = p->nx;
  a.ny = p->ny;
  a.rows = s.rows;
  a.dx = T(p->dx);
  a.dy = T(p->dy);
  a.cs = T(p->cs);
  a.cs2 = T(p->cs * p->cs);
  a.inv_cs = T(1.0 / p->cs);
  a.dfloor = T(p->dfloor);
  a.eps = T(std::max(1e-42, (double)std::numeric_limits<T>::min()));
  a.four_dx = T(4.0 * p->dx);
  a.four_dy = T(4.0 * p->dy);
  a.inv_dx = T(1.0 / p->dx);
  a.inv_dy = T(1.0 / p->dy);
  a.inv_four_dx = T(1.0 / (4.0 * p->dx));
  a.inv_four_dy = T(1.0 / (4.0 * p->dy));
  a.exact = exact;
  a.visc = T(p->viscosity);
  a.smag_on = p->smagorinsky_cs > 0.0;
  a.smag_c = T(p->smagorinsky_cs * p->smagorinsky_cs * p->dx * p->dy);
  a.drag_on = p->drag_rate > 0.0;
  a.drag = T(p->drag_rate);
  a.visc_terms = visc_terms ? 1 : 0;
  s.base = a;
  return s;
}

// Device memory a step uses beyond its input and output: the stage registers B, C and the
// RK4 saved register S (each one state), and the per-block partial sums of the kick.
// `timing`: bracket the stage launches with the context's HIP events.
template <typename T> struct StepScratch {
  T* B;
  T* C;
  T* S;
  double* partial;
  bool timing;
};

// Partial sums a step with the kick needs: the stage's per-block moments (at most
// kick_sums per block) and the second-moment pass (3 per aux block, at most 4096 blocks).
size_t kick_partials(size_t nblocks) { return std::max<size_t>(nblocks, 4096) * 7; }

// The context's own scratch for a step, allocated (grown) on demand.
template <typename T>
int ctx_scratch(jx_ctx* c, size_t state_bytes, size_t nblocks, bool kick, StepScratch<T>* out) {
  int rc = ensure_buffers(c, state_bytes);
  if (rc) return rc;
  if (kick) {
    rc = ensure_partial(c, kick_partials(nblocks));
    if (rc) return rc;
  }
  *out = StepScratch<T>{static_cast<T*>(c->buf[0]), static_cast<T*>(c->buf[1]),
                        static_cast<T*>(c->buf[2]), c->partial, c->timing};
  return JX_OK;
}

// Launch the RK stages of one step.  `first`: the saved register is the caller's input
// (floored on read).  The result lands in `A`.  `raw`: kick the final stage.  `want_dtkey`:
// the step's last kernel reduces the next CFL limit.  `kick_ts` (with `raw`): the kick runs
// inside the next step's timestep reduction (timestep_kernel<T, true> with these
// arguments), which reads the last stage's output from a scratch register and writes A.
template <typename T>
int launch_step(jx_ctx* c, const StepPlan<T>& s, const StepScratch<T>& w, const T* in, T* A,
                bool first, const T* raw, bool want_dtkey, const SgsHook<T>* hook = nullptr,
                const jx::TimestepArgs<T>* kick_ts = nullptr) {
  const bool fused_kick = raw && kick_ts;
  const jx_run_params* p = s.p;
  const RkTable tab = rk_table(p);
  const int ns = tab.n;
  T* B = w.B;
  T* C = w.C;
  T* S = w.S;
  const T* src = first ? in : A;
  const size_t nblocks = (size_t)s.grid.x * s.grid.y;
  // Register plan.  Stage outputs alternate B, C, B, ... and the last goes to A; a stage
  // never reads its stencil from the buffer it writes.  The saved register is the step's
  // input (floored on read when that is the caller's array) until a row with a delta
  // accumulates it into S (hydro.py:42-43); later stages read S.  Saved-register reads and
  // writes are pointwise, so they may share a buffer with the stage output.
  //   rk2: src->B, B|src->A          rk3: src->B, B|src->C, C|src->A
  //   rk4: src->B, B|src->C (S=src+d2 B), C|S->B (S+=d3 C), B|S->A (S+=0 B)
  // A one-stage table whose input is A itself writes B and copies it to A.
  bool saved_in_s = false;
  const T* prev = src;
  for (int k = 0; k < ns; ++k) {
    StageArgs<T> a = s.base;
    const RkRow& r = tab.row[k];
    a.gam0 = r.g0;
    a.gam1 = r.g1;
    a.beta = r.beta;
    a.delta = r.delta;
    a.has_delta = r.has_delta;
    a.blended = (r.g0 != 0.0 && r.g0 + r.g1 == 1.0) ? 1 : 0;
    const bool last = (k == ns - 1);
    a.stage1 = (k == 0 && plain_first_row(r)) ? 1 : 0;
    a.u_in = prev;
    if (last && fused_kick)
      a.u_out = (prev == B) ? C : B;   // the register the last stage does not read
    else
      a.u_out = (last && prev != A) ? A : ((k & 1) ? C : B);
    a.u_sv = a.stage1 ? nullptr : (saved_in_s ? S : src);
    a.floor_saved = (!a.stage1 && !saved_in_s && first) ? 1 : 0;
    a.s_out = r.has_delta ? S : nullptr;
    if (r.has_delta) saved_in_s = true;
    prev = a.u_out;
    a.raw = nullptr;
    a.partial = nullptr;
    a.dtkey = nullptr;
    if (last && raw) {
      a.raw = raw;
      a.partial = w.partial;
    } else if (last && want_dtkey) {
      a.dtkey = &c->ctrl->dtkey;
    }
    if (hook) {
      // recover_isothermal of the stage input, then the caller's stress (hydro.py:111-115).
      const size_t ncell = (size_t)p->nx * p->ny;
      hipLaunchKernelGGL(jx::recover_kernel<T>, aux_grid(ncell), dim3(jx::kAuxBlock), 0,
                         c->stream, a.u_in, static_cast<T*>(nullptr), hook->W, ncell, p->nvar,
                         T(p->dfloor));
      JX_HIP(c, hipGetLastError());
      if (hook->fn(hook->user, k, hook->W, hook->tau) != 0)
        return fail(c, JX_EINVAL, "the stress callback failed");
      a.tau = hook->tau;
    }
    std::pair<hipEvent_t, hipEvent_t>* ev = nullptr;
    if (w.timing) {
      if (c->events_used == c->events.size()) {
        std::pair<hipEvent_t, hipEvent_t> pr{};
        JX_HIP(c, hipEventCreate(&pr.first));
        JX_HIP(c, hipEventCreate(&pr.second));
        c->events.push_back(pr);
      }
      ev = &c->events[c->events_used++];
    }
    hipError_t e = StageLaunch<T>::go(p->reconstruct, p->rsolver, p->nvar, s.visc, a, s.grid,
                                      s.bx, c->stream, ev ? ev->first : nullptr,
                                      ev ? ev->second : nullptr);
    if (e != hipSuccess) return hip_fail(c, e, "stage kernel launch");
    if (last && a.u_out != A && !fused_kick)
      JX_HIP(c, hipMemcpyAsync(A, a.u_out, (size_t)p->nx * p->ny * p->nvar * sizeof(T),
                               hipMemcpyDeviceToDevice, c->stream));
  }
  const T* kick_src = prev;   // the last stage's output
  if (raw) {
    // Kick: forcing.py:137-149 then hydro.py:132, on the post-drag state in A.
    const size_t n = (size_t)p->nx * p->ny;
    const bool one_pass = s.visc && jx::kick_sums<T, true>() == 7;   // as the stage kernel
    hipLaunchKernelGGL(jx::reduce_kick_kernel, dim3(1), dim3(jx::kReduceBlock), 0, c->stream, c->ctrl,
                       w.partial, (int)nblocks, one_pass ? 2 : 0, (double)n);
    jx::KickArgs<T> ka{};
    ka.u = fused_kick ? const_cast<T*>(kick_src) : A;   // the moment pass reads it only
    ka.raw = raw;
    ka.nx = p->nx;
    ka.ny = p->ny;
    ka.nv = p->nvar;
    ka.dfloor = T(p->dfloor);
    ka.dx = T(p->dx);
    ka.dy = T(p->dy);
    ka.cs = T(p->cs);
    ka.ctrl = c->ctrl;
    ka.partial = w.partial;
    const dim3 g = aux_grid(n);
    if (!one_pass) {
      hipLaunchKernelGGL(jx::kick_second_moments_kernel<T>, g, dim3(jx::kAuxBlock), 0,
                         c->stream, ka);
      hipLaunchKernelGGL(jx::reduce_kick_kernel, dim3(1), dim3(jx::kReduceBlock), 0, c->stream, c->ctrl,
                         w.partial, (int)g.x, 1, (double)n);
    }
    if (fused_kick) {
      jx::TimestepArgs<T> t = *kick_ts;
      t.q = kick_src;
      t.primitive = 0;
      t.raw = raw;
      t.out = A;
      t.waves = jx::kTsKickWaves;
      hipLaunchKernelGGL((jx::timestep_kernel<T, true>), jx::ts_grid(p->nx, p->ny, t.waves),
                         dim3(jx::kAuxBlock), 0, c->stream, t);
    } else {
      ka.dtkey = want_dtkey ? &c->ctrl->dtkey : nullptr;
      hipLaunchKernelGGL(jx::kick_apply_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream, ka);
    }
  }
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

template <typename T>
jx::TimestepArgs<T> timestep_args(jx_ctx* c, const jx_run_params* p, const T* q, int primitive,
                                  const Ctrl* ctrl) {
  jx::TimestepArgs<T> t{};
  t.q = q;
  t.nx = p->nx;
  t.ny = p->ny;
  t.nv = p->nvar;
  t.primitive = primitive;
  t.smag_on = p->smagorinsky_cs > 0.0;
  t.dx = T(p->dx);
  t.dy = T(p->dy);
  t.cs = T(p->cs);
  t.dfloor = T(p->dfloor);
  t.four_dx = T(4.0 * p->dx);
  t.four_dy = T(4.0 * p->dy);
  t.smag_c = T(p->smagorinsky_cs * p->smagorinsky_cs * p->dx * p->dy);
  t.ctrl = ctrl;
  t.dtkey = &c->ctrl->dtkey;
  t.nukey = &c->ctrl->nukey;
  t.waves = jx::kTsWaves;
  return t;
}

template <typename T>
int step_impl(jx_ctx* c, const jx_run_params* p, const T* in, T* out, double dt, const T* raw,
              double dedt, const SgsHook<T>* hook = nullptr) {
  const size_t bytes = (size_t)p->nx * p->ny * p->nvar * sizeof(T);
  StepPlan<T> s = make_plan<T>(c, p, hook != nullptr);
  StepScratch<T> w{};
  int rc = ctx_scratch<T>(c, bytes, (size_t)s.grid.x * s.grid.y, raw != nullptr, &w);
  if (rc) return rc;
  hipLaunchKernelGGL(jx::set_step_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl, dt, dedt);
  return launch_step<T>(c, s, w, in, out, true, raw, false, hook);
}

// One step on the static scratch: the same plan and kernels as step_impl, with the stage
// registers and partial sums carved out of `base` (this device's address of `scratch`).
// Nothing there needs to start at zero: every stage writes its whole output register
// before a later stage reads it, S is written from the step's input by the first row with
// a delta, and each partial sum is written by its block before the reduction reads it.
template <typename T>
int step_noalloc_impl(jx_ctx* c, const jx_run_params* p, const T* in, T* out, double dt,
                      const T* raw, double dedt, unsigned char* base) {
  const size_t bytes = (size_t)p->nx * p->ny * p->nvar * sizeof(T);
  const size_t reg = (bytes + 255) & ~(size_t)255;
  const RkTable tab = rk_table(p);
  bool delta = false;
  for (int k = 0; k < tab.n; ++k) delta = delta || tab.row[k].has_delta;
  const size_t nreg = delta ? 3 : 2;
  if (kNoAllocPartialBytes + nreg * reg > kNoAllocScratchBytes)
    return fail(c, JX_EINVAL,
                "jx_isothermal_step_noalloc: this state needs " + std::to_string(nreg) +
                    " stage registers of " + std::to_string(bytes) + " bytes; the static "
                    "scratch holds " + std::to_string(kNoAllocScratchBytes -
                                                      kNoAllocPartialBytes) +
                    " bytes of registers (use jx_isothermal_step for larger states)");
  StepPlan<T> s = make_plan<T>(c, p, false);
  const size_t nblocks = (size_t)s.grid.x * s.grid.y;
  if (raw && kick_partials(nblocks) * sizeof(double) > kNoAllocPartialBytes)
    return fail(c, JX_EINVAL, "jx_isothermal_step_noalloc: the kick needs partial sums for " +
                                  std::to_string(nblocks) + " blocks; the static scratch "
                                  "holds them for at most " +
                                  std::to_string(kNoAllocPartialBytes / (7 * sizeof(double))));
  unsigned char* regs = base + kNoAllocPartialBytes;
  StepScratch<T> w{reinterpret_cast<T*>(regs), reinterpret_cast<T*>(regs + reg),
                   delta ? reinterpret_cast<T*>(regs + 2 * reg) : nullptr,
                   reinterpret_cast<double*>(base), false};
  hipLaunchKernelGGL(jx::set_step_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl, dt, dedt);
  return launch_step<T>(c, s, w, in, out, true, raw, false, nullptr);
}

// FP32 forced runs with Smagorinsky (their final stage takes the kick moments in one pass
// and the kick runs inside the timestep reduction, the only two readers of the force)
// evaluate the force where it is used when the basis has at most kMaxForceGroups mode
// groups: the plan's stages and `*fe` get the per-row group sums (c->force_pq).  Reads the
// group count back once (the groups are found on the device by prepare_render).
template <typename T>
int prepare_force_eval(jx_ctx* c, const jx_run_params* p, const jx_forcing* f, bool smag,
                       StepPlan<T>* s, jx::ForceEval<T>* fe) {
  *fe = jx::ForceEval<T>{};
  if constexpr (std::is_same<T, float>::value) {
    if (!smag || !c->fused_forcing) return JX_OK;
    int* g = reinterpret_cast<int*>(c->host_sums);
    JX_HIP(c, hipMemcpyAsync(g, c->groups, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    JX_HIP(c, hipStreamSynchronize(c->stream));
    if (*g < 1 || *g > jx::kMaxForceGroups) return JX_OK;
    int rc = ensure(c, &c->force_pq, &c->force_pq_bytes, force_pq_bytes(p->ny));
    if (rc) return rc;
    fe->pq = static_cast<const float*>(c->force_pq);
    fe->cx = static_cast<const float*>(f->cos_x);
    fe->sx = static_cast<const float*>(f->sin_x);
    fe->grp = static_cast<const int*>(c->groups);
    fe->n_modes = f->n_modes;
    s->base.force = *fe;
  }
  return JX_OK;
}

template <typename T>
int evolve_impl(jx_ctx* c, const jx_run_params* p, const T* in, T* out, const jx_forcing* f,
                double* t_out, double* dt_out, int* cyc_out, const SgsHook<T>* hook = nullptr,
                const jx_loop_state* start = nullptr, jx_loop_state* end = nullptr) {
  const size_t n = (size_t)p->nx * p->ny;
  const size_t bytes = n * p->nvar * sizeof(T);
  const bool forced = f != nullptr;
  StepPlan<T> s = make_plan<T>(c, p, hook != nullptr);
  StepScratch<T> w{};
  int rc = ctx_scratch<T>(c, bytes, (size_t)s.grid.x * s.grid.y, forced, &w);
  if (rc) return rc;
  if (forced) {
    rc = ensure(c, &c->raw, &c->raw_bytes, n * 2 * sizeof(T));
    if (rc) return rc;
  }
  hipLaunchKernelGGL(jx::init_ctrl_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl);
  const jx_loop_state s0 = start ? *start : jx_loop_state{0.0, INFINITY, 0, -1};
  if (start)
    hipLaunchKernelGGL(jx::set_loop_state_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl,
                       s0.time, s0.prev_dt, s0.cycle, forced ? s0.update : -1);
  if (!(s0.time < p->tlim && s0.cycle < p->nlim)) {
    // No step to take (run.py:36-38): the state passes through.
    JX_HIP(c, hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, c->stream));
    JX_HIP(c, hipStreamSynchronize(c->stream));
    *t_out = s0.time;
    *dt_out = s0.prev_dt;
    *cyc_out = s0.cycle;
    if (end) *end = s0;
    return JX_OK;
  }
  const bool smag = p->smagorinsky_cs > 0.0;
  jx::ForceEval<T> ts_force{};   // pq set: the force is evaluated where it is used
  jx::LoopArgs la{};
  la.tlim = p->tlim;
  la.cfl = p->cfl;
  la.viscosity = p->viscosity;
  la.drag_rate = p->drag_rate;
  la.h = std::min(p->dx, p->dy);
  la.dt_update = p->dt_update;
  la.tdriv_start = p->tdriv_start;
  la.dedt = forced ? p->dedt : 0.0;
  la.nlim = p->nlim;
  la.smag_on = smag;
  la.has_forcing = forced;
  la.n_updates = forced ? f->n_updates : 0;
  jx::RenderArgs<T> ra{};
  if (forced) {
    ra.raw = static_cast<T*>(c->raw);
    ra.table = static_cast<const T*>(f->coefficients);
    ra.cx = static_cast<const T*>(f->cos_x);
    ra.sx = static_cast<const T*>(f->sin_x);
    ra.cy = static_cast<const T*>(f->cos_y);
    ra.sy = static_cast<const T*>(f->sin_y);
    ra.nx = p->nx;
    ra.ny = p->ny;
    ra.n_modes = f->n_modes;
    ra.n_updates = f->n_updates;
    ra.ctrl = c->ctrl;
    rc = prepare_render<T>(c, ra);
    if (rc) return rc;
    rc = prepare_force_eval<T>(c, p, f, smag, &s, &ts_force);
    if (rc) return rc;
    if (ts_force.pq)
      JX_HIP(c, hipMemsetAsync(c->force_pq, 0, force_pq_bytes(p->ny), c->stream));
    else
      JX_HIP(c, hipMemsetAsync(c->raw, 0, n * 2 * sizeof(T), c->stream));
    if (s0.update >= 0) {
      // Resume: the raw force of the OU row in effect (forcing.py:111-122).
      jx::RenderArgs<T> r0 = ra;
      r0.ctrl = nullptr;
      r0.row = s0.update;
      launch_render<T>(c, r0, ts_force.pq != nullptr);
    }
  }
  const bool eval_force = ts_force.pq != nullptr;
  const T* raw = forced ? static_cast<const T*>(c->raw) : nullptr;
  // Initial timestep from the caller's state (run.py:42-46).  Each step opens with the
  // one-thread begin_step kernel (commit the last step, test the loop condition, pick dt).
  hipLaunchKernelGGL(jx::timestep_kernel<T>, jx::ts_grid(p->nx, p->ny), dim3(jx::kAuxBlock), 0,
                     c->stream, timestep_args<T>(c, p, in, 0, c->ctrl));
  jx::TimestepArgs<T> ts = timestep_args<T>(c, p, out, 0, c->ctrl);
  ts.force = ts_force;
  // `open`: launch the step's begin_step (false for the first step after a chunk boundary,
  // which the boundary's begin_step has opened already).
  auto one_step = [&](bool first, bool open) -> int {
    if (open)
      hipLaunchKernelGGL(jx::begin_step_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl, la);
    if (forced) launch_render<T>(c, ra, eval_force);
    // Forced runs with Smagorinsky kick inside the next timestep reduction.
    const bool fuse = forced && smag;
    int r = launch_step<T>(c, s, w, in, out, first, raw, !smag, hook, fuse ? &ts : nullptr);
    if (r) return r;
    if (smag && !fuse)
      hipLaunchKernelGGL(jx::timestep_kernel<T>, jx::ts_grid(p->nx, p->ny), dim3(jx::kAuxBlock), 0,
                         c->stream, ts);
    return JX_OK;
  };
  // The first chunk goes out before dt is read back: up to 32 steps (kernels launched past
  // the end of the loop exit at once).  With a stress callback every step is its own chunk:
  // the callback is host code, and steps past the end must not call it.
  long long chunk = hook ? 1 : std::min<long long>(32, (long long)p->nlim - s0.cycle);
  bool first = true, opened = false;
  for (;;) {
    for (long long k = 0; k < chunk; ++k) {
      rc = one_step(first, !(k == 0 && opened));
      if (rc) return rc;
      first = false;
    }
    opened = true;
    // Commit the last launched step (and open the next) before reading the loop state.
    hipLaunchKernelGGL(jx::begin_step_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl, la);
    JX_HIP(c, hipMemcpyAsync(c->host_ctrl, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost,
                             c->stream));
    JX_HIP(c, hipStreamSynchronize(c->stream));
    const Ctrl h = *c->host_ctrl;
    if (h.done) {
      *t_out = h.time;
      *dt_out = h.prev_dt;
      *cyc_out = h.cycle;
      if (end) *end = jx_loop_state{h.time, h.prev_dt, h.cycle, h.update};
      return JX_OK;
    }
    // The next step is open: launch it and a chunk more.  Remaining steps are estimated
    // from the time left; dt can shrink, so loop.
    double est = (p->tlim - h.time) / h.dt;
    long long want = std::isfinite(est) ? (long long)std::ceil(est) : 1;
    want = std::min<long long>(want, (long long)p->nlim - h.cycle);
    chunk = hook ? 1 : std::max<long long>(1, std::min<long long>(want, 512));
  }
}

template <typename T>
int normalize_impl(jx_ctx* c, int ny, int nx, int nvar, const T* U, const T* raw, double dt,
                          double dedt, T* acc) {
  const size_t n = (size_t)nx * ny;
  const dim3 g = aux_grid(n);
  int rc = ensure_partial(c, (size_t)g.x * 3);
  if (rc) return rc;
  hipLaunchKernelGGL(jx::set_step_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl, dt, dedt);
  jx::KickArgs<T> ka{};
  ka.u = const_cast<T*>(U);
  ka.raw = raw;
  ka.nx = nx;
  ka.ny = ny;
  ka.nv = nvar;
  ka.ctrl = c->ctrl;
  ka.partial = c->partial;
  hipLaunchKernelGGL(jx::kick_first_moments_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream, ka);
  hipLaunchKernelGGL(jx::reduce_kick_kernel, dim3(1), dim3(jx::kReduceBlock), 0, c->stream, c->ctrl,
                     c->partial, (int)g.x, 0, (double)n);
  hipLaunchKernelGGL(jx::kick_second_moments_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream, ka);
  hipLaunchKernelGGL(jx::reduce_kick_kernel, dim3(1), dim3(jx::kReduceBlock), 0, c->stream, c->ctrl,
                     c->partial, (int)g.x, 1, (double)n);
  hipLaunchKernelGGL(jx::kick_acceleration_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream,
                     (const Ctrl*)c->ctrl, raw, acc, n);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}


// ---- unfused stage operators (traced and learned-stress stage paths) ------------------

template <typename T>
int face_states_impl(jx_ctx* c, int ny, int nx, int nv, const T* W, int axis, int recon,
                     double dfloor, T* L, T* R) {
  const dim3 g = aux_grid((size_t)nx * ny);
  const T eps = T(std::max(1e-42, (double)std::numeric_limits<T>::min()));
#define JX_FACES(RC)                                                                      \
  hipLaunchKernelGGL((jx::face_states_kernel<T, RC>), g, dim3(jx::kAuxBlock), 0, c->stream, \
                     W, L, R, nx, ny, nv, axis, T(dfloor), eps)
  switch (recon) {
    case jx::RECON_DC: JX_FACES(jx::RECON_DC); break;
    case jx::RECON_PLM: JX_FACES(jx::RECON_PLM); break;
    case jx::RECON_PPM4: JX_FACES(jx::RECON_PPM4); break;
    case jx::RECON_PPMX: JX_FACES(jx::RECON_PPMX); break;
    case jx::RECON_WENOZ: JX_FACES(jx::RECON_WENOZ); break;
    default: return fail(c, JX_EKEY, "unknown reconstruct id");
  }
#undef JX_FACES
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

template <typename T>
int riemann_impl(jx_ctx* c, size_t n, int nv, const T* L, const T* R, double cs, int axis,
                 int rs, T* F) {
  const dim3 g = aux_grid(n);
  const jx::Gas<T> gas{T(cs), T(cs * cs), T(1.0 / cs)};
#define JX_FLUX(RS_)                                                                       \
  hipLaunchKernelGGL((jx::riemann_faces_kernel<T, RS_>), g, dim3(jx::kAuxBlock), 0,         \
                     c->stream, L, R, F, n, nv, axis, gas)
  switch (rs) {
    case jx::RS_LLF: JX_FLUX(jx::RS_LLF); break;
    case jx::RS_HLLE: JX_FLUX(jx::RS_HLLE); break;
    case jx::RS_ROE: JX_FLUX(jx::RS_ROE); break;
    default: return fail(c, JX_EKEY, "unknown rsolver id");
  }
#undef JX_FLUX
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

template <typename T>
int viscous_impl(jx_ctx* c, int ny, int nx, int nv, const T* W, double dx, double dy,
                 double nu, double smag_cs, int axis, T* out) {
  const dim3 g = aux_grid((size_t)nx * ny);
  const jx::Visc<T> vp{T(nu), T(smag_cs * smag_cs * dx * dy), smag_cs > 0.0};
  const jx::Over<T, false> odx{T(dx), T(1.0 / dx)}, ody{T(dy), T(1.0 / dy)};
  const jx::Over<T, false> o4dx{T(4.0 * dx), T(1.0 / (4.0 * dx))};
  const jx::Over<T, false> o4dy{T(4.0 * dy), T(1.0 / (4.0 * dy))};
  hipLaunchKernelGGL(jx::viscous_faces_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream, W, out,
                     nx, ny, nv, axis, odx, ody, o4dx, o4dy, vp);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

template <typename T>
int gradients_impl(jx_ctx* c, int ny, int nx, int nv, const T* W, double dx, double dy,
                   int axis, T* dvdx, T* dvdy) {
  const dim3 g = aux_grid((size_t)nx * ny);
  const jx::Over<T, false> odx{T(dx), T(1.0 / dx)}, ody{T(dy), T(1.0 / dy)};
  const jx::Over<T, false> o4dx{T(4.0 * dx), T(1.0 / (4.0 * dx))};
  const jx::Over<T, false> o4dy{T(4.0 * dy), T(1.0 / (4.0 * dy))};
  hipLaunchKernelGGL(jx::face_gradients_kernel<T>, g, dim3(jx::kAuxBlock), 0, c->stream, W,
                     dvdx, dvdy, nx, ny, nv, axis, odx, ody, o4dx, o4dy);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

// evolve_isothermal with an optional stress callback and loop state (shared checks).
int evolve_entry(jx_ctx* c, const jx_run_params* p, const void* U_in, void* U_out,
                 const jx_forcing* f, void* W, void* tau, jx_stress_fn fn, void* user,
                 double* time_out, double* dt_out, int* cycle_out, const jx_loop_state* start,
                 jx_loop_state* end) {
  if (!c || !time_out || !dt_out || !cycle_out) return JX_EINVAL;
  if (fn && (!W || !tau)) return fail(c, JX_EINVAL, "a stress callback needs W and tau buffers");
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!std::isfinite(p->dedt) || p->dedt < 0.0)
    return fail(c, JX_EINVAL, "dedt must be finite and nonnegative");
  if (f && (!(p->dt_update > 0.0) || !std::isfinite(p->dt_update) ||
            !std::isfinite(p->tdriv_start)))
    return fail(c, JX_EINVAL, "dt_update must be finite and positive and tdriv_start finite");
  if (f && (!f->coefficients || !f->cos_x || !f->sin_x || !f->cos_y || !f->sin_y ||
            f->n_modes < 1 || f->n_updates < 1))
    return fail(c, JX_EINVAL, "incomplete forcing tables");
  if (!U_in || !U_out || U_in == U_out)
    return fail(c, JX_EINVAL, "U_in and U_out must be distinct device arrays");
  if (p->dtype == 0) {
    SgsHook<double> h{static_cast<double*>(W), static_cast<double*>(tau), fn, user};
    return evolve_impl<double>(c, p, static_cast<const double*>(U_in),
                               static_cast<double*>(U_out), f, time_out, dt_out, cycle_out,
                               fn ? &h : nullptr, start, end);
  }
  SgsHook<float> h{static_cast<float*>(W), static_cast<float*>(tau), fn, user};
  return evolve_impl<float>(c, p, static_cast<const float*>(U_in), static_cast<float*>(U_out), f,
                            time_out, dt_out, cycle_out, fn ? &h : nullptr, start, end);
}

}  // namespace

extern "C" {

int jx_ctx_create(jx_ctx** out, int device, void* stream) {
  if (!out) return JX_EINVAL;
  *out = nullptr;
  jx_ctx* c = new jx_ctx();
  c->device = device;
  c->stream = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(&c->ctrl, sizeof(Ctrl));
  if (e == hipSuccess) e = hipHostMalloc(&c->host_ctrl, sizeof(Ctrl), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&c->sums, 16 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&c->host_sums, 16 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) scratch_setup(device);
  if (e != hipSuccess) {
    std::fprintf(stderr, "jx_ctx_create: %s\n", hipGetErrorString(e));
    jx_ctx_destroy(c);
    return JX_EHIP;
  }
  *out = c;
  return JX_OK;
}

void jx_ctx_destroy(jx_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  else (void)hipDeviceSynchronize();
  for (auto& b : c->buf)
    if (b) (void)hipFree(b);
  if (c->raw) (void)hipFree(c->raw);
  if (c->groups) (void)hipFree(c->groups);
  if (c->force_pq) (void)hipFree(c->force_pq);
  if (c->partial) (void)hipFree(c->partial);
  if (c->ctrl) (void)hipFree(c->ctrl);
  if (c->sums) (void)hipFree(c->sums);
  if (c->host_ctrl) (void)hipHostFree(c->host_ctrl);
  if (c->host_sums) (void)hipHostFree(c->host_sums);
  for (auto& pr : c->events) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  delete c;
}

const char* jx_last_error(const jx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int jx_set_tiles(jx_ctx* c, int block_x, int rows) {
  if (!c) return JX_EINVAL;
  if (block_x != 0 && (block_x < 64 || block_x > jx::kMaxBlock || block_x % 64 != 0))
    return fail(c, JX_EINVAL, "block_x must be 0 or a multiple of 64 in [64, 256]");
  if (rows < 0) return fail(c, JX_EINVAL, "rows must be >= 0");
  c->block_x = block_x;
  c->rows = rows;
  return JX_OK;
}

int jx_get_tiles(jx_ctx* c, const jx_run_params* p, int* block_x, int* rows) {
  if (!c || !p || !block_x || !rows) return JX_EINVAL;
  int rc = check_params(c, p);
  if (rc) return rc;
  const bool visc = p->viscosity > 0.0 || p->smagorinsky_cs > 0.0;
  const bool exact = is_pow2(p->dx) && is_pow2(p->dy) && is_pow2(p->cs);
  if (p->dtype == 1) pick_tiles<float>(c, p, visc, exact, block_x, rows);
  else pick_tiles<double>(c, p, visc, exact, block_x, rows);
  return JX_OK;
}

int jx_set_fused_forcing(jx_ctx* c, int on) {
  if (!c) return JX_EINVAL;
  c->fused_forcing = on != 0;
  return JX_OK;
}

int jx_set_kernel_timing(jx_ctx* c, int on) {
  if (!c) return JX_EINVAL;
  c->timing = on != 0;
  c->events_used = 0;
  return JX_OK;
}

int jx_kernel_timing(jx_ctx* c, int* launches, double* total_ms) {
  if (!c || !launches || !total_ms) return JX_EINVAL;
  JX_HIP(c, hipStreamSynchronize(c->stream));
  double total = 0.0;
  int n = 0;
  for (size_t k = 0; k < c->events_used; ++k) {
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, c->events[k].first, c->events[k].second);
    if (e != hipSuccess) return hip_fail(c, e, "hipEventElapsedTime");
    total += ms;
    ++n;
  }
  c->events_used = 0;
  *launches = n;
  *total_ms = total;
  return JX_OK;
}

int jx_isothermal_step(jx_ctx* c, const jx_run_params* p, const void* U_in, void* U_out,
                       double dt, const void* raw_force, double dedt) {
  return jx_isothermal_step_sgs(c, p, U_in, U_out, dt, raw_force, dedt, nullptr, nullptr,
                                nullptr, nullptr);
}

int jx_isothermal_step_noalloc(jx_ctx* c, const jx_run_params* p, const void* U_in,
                               void* U_out, double dt, const void* raw_force, double dedt) {
  if (!c) return JX_EINVAL;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!U_in || !U_out || U_in == U_out)
    return fail(c, JX_EINVAL, "U_in and U_out must be distinct device arrays");
  hipError_t status = hipSuccess;
  void* base = scratch_addr(c->device, &status);
  if (!base)
    return fail(c, JX_EHIP, std::string("jx_isothermal_step_noalloc: no static scratch on "
                                        "device ") + std::to_string(c->device) + ": " +
                                hipGetErrorString(status));
  unsigned char* b = static_cast<unsigned char*>(base);
  if (p->dtype == 0)
    return step_noalloc_impl<double>(c, p, static_cast<const double*>(U_in),
                                     static_cast<double*>(U_out), dt,
                                     static_cast<const double*>(raw_force), dedt, b);
  return step_noalloc_impl<float>(c, p, static_cast<const float*>(U_in),
                                  static_cast<float*>(U_out), dt,
                                  static_cast<const float*>(raw_force), dedt, b);
}

int jx_isothermal_step_sgs(jx_ctx* c, const jx_run_params* p, const void* U_in, void* U_out,
                           double dt, const void* raw_force, double dedt, void* W, void* tau,
                           jx_stress_fn fn, void* user) {
  if (!c) return JX_EINVAL;
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!U_in || !U_out || U_in == U_out)
    return fail(c, JX_EINVAL, "U_in and U_out must be distinct device arrays");
  if (fn && (!W || !tau)) return fail(c, JX_EINVAL, "a stress callback needs W and tau buffers");
  if (p->dtype == 0) {
    SgsHook<double> h{static_cast<double*>(W), static_cast<double*>(tau), fn, user};
    return step_impl<double>(c, p, static_cast<const double*>(U_in), static_cast<double*>(U_out),
                             dt, static_cast<const double*>(raw_force), dedt, fn ? &h : nullptr);
  }
  SgsHook<float> h{static_cast<float*>(W), static_cast<float*>(tau), fn, user};
  return step_impl<float>(c, p, static_cast<const float*>(U_in), static_cast<float*>(U_out), dt,
                          static_cast<const float*>(raw_force), dedt, fn ? &h : nullptr);
}

int jx_isothermal_timestep(jx_ctx* c, const jx_run_params* p, const void* W, double previous_dt,
                           double time, double tlim, double* dt_out) {
  if (!c || !dt_out) return JX_EINVAL;
  if (!p || !W) return fail(c, JX_EINVAL, "params and W must not be NULL");
  if (p->nx < 1 || p->ny < 1 || p->nvar < 3) return fail(c, JX_EINVAL, "bad shape");
  const dim3 grid = jx::ts_grid(p->nx, p->ny);
  hipLaunchKernelGGL(jx::init_ctrl_kernel, dim3(1), dim3(1), 0, c->stream, c->ctrl);
  if (p->dtype == 0)
    hipLaunchKernelGGL(jx::timestep_kernel<double>, grid, dim3(jx::kAuxBlock), 0, c->stream,
                       timestep_args<double>(c, p, static_cast<const double*>(W), 1, nullptr));
  else
    hipLaunchKernelGGL(jx::timestep_kernel<float>, grid, dim3(jx::kAuxBlock), 0, c->stream,
                       timestep_args<float>(c, p, static_cast<const float*>(W), 1, nullptr));
  JX_HIP(c, hipGetLastError());
  JX_HIP(c, hipMemcpyAsync(c->host_ctrl, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost,
                           c->stream));
  JX_HIP(c, hipStreamSynchronize(c->stream));
  const Ctrl h = *c->host_ctrl;
  *dt_out = jx::finish_timestep(jx::value_of_key(h.dtkey), jx::value_of_max_key(h.nukey),
                                p->smagorinsky_cs > 0.0, p->viscosity, p->drag_rate,
                                std::min(p->dx, p->dy), p->cfl, previous_dt, time, tlim);
  return JX_OK;
}

int jx_evolve_isothermal(jx_ctx* c, const jx_run_params* p, const void* U_in, void* U_out,
                         const jx_forcing* f, double* time_out, double* dt_out, int* cycle_out) {
  return jx_evolve_isothermal_sgs(c, p, U_in, U_out, f, nullptr, nullptr, nullptr, nullptr,
                                  time_out, dt_out, cycle_out);
}

int jx_evolve_isothermal_resume(jx_ctx* c, const jx_run_params* p, const void* U_in,
                                void* U_out, const jx_forcing* f, const jx_loop_state* start,
                                jx_loop_state* end) {
  if (!c) return JX_EINVAL;
  if (start && (!std::isfinite(start->time) || start->cycle < 0 || start->update < -1 ||
                !(start->prev_dt > 0.0)))
    return fail(c, JX_EINVAL, "bad loop state");
  double t, dt;
  int cyc;
  return evolve_entry(c, p, U_in, U_out, f, nullptr, nullptr, nullptr, nullptr, &t, &dt, &cyc,
                      start, end);
}

int jx_evolve_isothermal_sgs(jx_ctx* c, const jx_run_params* p, const void* U_in, void* U_out,
                             const jx_forcing* f, void* W, void* tau, jx_stress_fn fn,
                             void* user, double* time_out, double* dt_out, int* cycle_out) {
  return evolve_entry(c, p, U_in, U_out, f, W, tau, fn, user, time_out, dt_out, cycle_out,
                      nullptr, nullptr);
}

int jx_recover_isothermal(jx_ctx* c, int dtype, long long ncells, int nvar, const void* U,
                          double dfloor, void* U_out, void* W_out) {
  if (!c) return JX_EINVAL;
  if (ncells < 0 || nvar < 1 || !U) return fail(c, JX_EINVAL, "bad recover arguments");
  const dim3 g = aux_grid((size_t)ncells);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::recover_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(U), static_cast<double*>(U_out),
                       static_cast<double*>(W_out), (size_t)ncells, nvar, dfloor);
  else
    hipLaunchKernelGGL(jx::recover_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(U), static_cast<float*>(U_out),
                       static_cast<float*>(W_out), (size_t)ncells, nvar, (float)dfloor);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_rk_stage(jx_ctx* c, int dtype, long long n, const void* U, const void* U_saved,
                const void* div_flux, double dt, double gam0, double gam1, double beta,
                double delta, int has_delta, void* U_next, void* U_saved_out) {
  if (!c) return JX_EINVAL;
  if (n < 0 || !U || !U_saved || !div_flux || !U_next)
    return fail(c, JX_EINVAL, "bad rk_stage arguments");
  jx::StageCoef k{gam0, gam1, beta, delta, has_delta, (gam0 != 0.0 && gam0 + gam1 == 1.0)};
  const dim3 g = aux_grid((size_t)n);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::rk_stage_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(U), static_cast<const double*>(U_saved),
                       static_cast<const double*>(div_flux), static_cast<double*>(U_next),
                       static_cast<double*>(U_saved_out), (size_t)n, dt, k);
  else
    hipLaunchKernelGGL(jx::rk_stage_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(U), static_cast<const float*>(U_saved),
                       static_cast<const float*>(div_flux), static_cast<float*>(U_next),
                       static_cast<float*>(U_saved_out), (size_t)n, dt, k);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_flux_divergence(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* fx,
                       const void* fy, double dx, double dy, void* out) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 1 || !fx || !fy || !out)
    return fail(c, JX_EINVAL, "bad flux_divergence arguments");
  const dim3 g = aux_grid((size_t)nx * ny * nvar);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::flux_divergence_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(fx), static_cast<const double*>(fy),
                       static_cast<double*>(out), nx, ny, nvar, dx, dy);
  else
    hipLaunchKernelGGL(jx::flux_divergence_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(fx), static_cast<const float*>(fy),
                       static_cast<float*>(out), nx, ny, nvar, (float)dx, (float)dy);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_stress_face_fluxes(jx_ctx* c, int dtype, int ny, int nx, const void* tau, void* fx,
                          void* fy) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || !tau || !fx || !fy)
    return fail(c, JX_EINVAL, "Expected stress with shape (ny, nx, 3)");
  const dim3 g = aux_grid((size_t)nx * ny);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::stress_faces_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(tau), static_cast<double*>(fx),
                       static_cast<double*>(fy), nx, ny);
  else
    hipLaunchKernelGGL(jx::stress_faces_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(tau), static_cast<float*>(fx),
                       static_cast<float*>(fy), nx, ny);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_render_force(jx_ctx* c, int dtype, int ny, int nx, int n_modes, const void* amplitudes,
                    const void* cos_x, const void* sin_x, const void* cos_y, const void* sin_y,
                    void* raw_out) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || n_modes < 0 || !raw_out || (n_modes > 0 && !amplitudes))
    return fail(c, JX_EINVAL, "bad render_force arguments");
  const dim3 g = render_grid(nx, ny);
  if (dtype == 0) {
    jx::RenderArgs<double> a{static_cast<double*>(raw_out), static_cast<const double*>(amplitudes),
                             static_cast<const double*>(cos_x), static_cast<const double*>(sin_x),
                             static_cast<const double*>(cos_y), static_cast<const double*>(sin_y),
                             nx, ny, n_modes, 1, nullptr, 0};
    hipLaunchKernelGGL(jx::render_force_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream, a);
  } else {
    jx::RenderArgs<float> a{static_cast<float*>(raw_out), static_cast<const float*>(amplitudes),
                            static_cast<const float*>(cos_x), static_cast<const float*>(sin_x),
                            static_cast<const float*>(cos_y), static_cast<const float*>(sin_y),
                            nx, ny, n_modes, 1, nullptr, 0};
    int rc = prepare_render<float>(c, a);
    if (rc) return rc;
    launch_render<float>(c, a);
  }
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_normalize_force(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* U,
                       const void* raw_force, double dt, double dedt, void* acc_out) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 3 || !U || !raw_force || !acc_out)
    return fail(c, JX_EINVAL, "bad normalize_force arguments");
  if (dtype == 0)
    return normalize_impl<double>(c, ny, nx, nvar, static_cast<const double*>(U),
                                  static_cast<const double*>(raw_force), dt, dedt,
                                  static_cast<double*>(acc_out));
  return normalize_impl<float>(c, ny, nx, nvar, static_cast<const float*>(U),
                               static_cast<const float*>(raw_force), dt, dedt,
                               static_cast<float*>(acc_out));
}

int jx_conserved_sums(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* U,
                      double* sums_out) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 1 || nvar > 8 || !U || !sums_out)
    return fail(c, JX_EINVAL, "bad conserved_sums arguments");
  const size_t n = (size_t)nx * ny;
  const dim3 g = aux_grid(n);
  int rc = ensure_partial(c, (size_t)g.x * 8);
  if (rc) return rc;
  if (dtype == 0)
    hipLaunchKernelGGL(jx::box_sums_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(U), n, nvar, c->partial);
  else
    hipLaunchKernelGGL(jx::box_sums_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(U), n, nvar, c->partial);
  hipLaunchKernelGGL(jx::box_sums_finish_kernel, dim3(1), dim3(64), 0, c->stream, c->partial,
                     (int)g.x, nvar, c->sums);
  JX_HIP(c, hipGetLastError());
  JX_HIP(c, hipMemcpyAsync(c->host_sums, c->sums, nvar * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  JX_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(sums_out, c->host_sums, nvar * sizeof(double));
  return JX_OK;
}

// ---- unfused stage operators ------------------------------------------------------------

int jx_reconstruct_faces(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* W,
                         int axis, int reconstruct, double dfloor, void* left, void* right) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 1 || (axis != 0 && axis != 1) || !W || !left || !right)
    return fail(c, JX_EINVAL, "bad reconstruct_faces arguments");
  if (dtype == 0)
    return face_states_impl<double>(c, ny, nx, nvar, static_cast<const double*>(W), axis,
                                    reconstruct, dfloor, static_cast<double*>(left),
                                    static_cast<double*>(right));
  return face_states_impl<float>(c, ny, nx, nvar, static_cast<const float*>(W), axis,
                                 reconstruct, dfloor, static_cast<float*>(left),
                                 static_cast<float*>(right));
}

int jx_isothermal_flux(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* left,
                       const void* right, double cs, int axis, int rsolver, void* flux) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 3 || (axis != 0 && axis != 1) || !left || !right || !flux ||
      !(cs > 0.0))
    return fail(c, JX_EINVAL, "bad isothermal_flux arguments");
  const size_t n = (size_t)nx * ny;
  if (dtype == 0)
    return riemann_impl<double>(c, n, nvar, static_cast<const double*>(left),
                                static_cast<const double*>(right), cs, axis, rsolver,
                                static_cast<double*>(flux));
  return riemann_impl<float>(c, n, nvar, static_cast<const float*>(left),
                             static_cast<const float*>(right), cs, axis, rsolver,
                             static_cast<float*>(flux));
}

int jx_viscous_momentum_flux(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* W,
                             double dx, double dy, double viscosity, double smagorinsky_cs,
                             int axis, void* out) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 3 || (axis != 0 && axis != 1) || !W || !out)
    return fail(c, JX_EINVAL, "bad viscous_momentum_flux arguments");
  if (dtype == 0)
    return viscous_impl<double>(c, ny, nx, nvar, static_cast<const double*>(W), dx, dy,
                                viscosity, smagorinsky_cs, axis, static_cast<double*>(out));
  return viscous_impl<float>(c, ny, nx, nvar, static_cast<const float*>(W), dx, dy, viscosity,
                             smagorinsky_cs, axis, static_cast<float*>(out));
}


int jx_face_velocity_gradients(jx_ctx* c, int dtype, int ny, int nx, int nvar, const void* W,
                               double dx, double dy, int axis, void* dv_dx, void* dv_dy) {
  if (!c) return JX_EINVAL;
  if (ny < 1 || nx < 1 || nvar < 3 || (axis != 0 && axis != 1) || !W || !dv_dx || !dv_dy)
    return fail(c, JX_EINVAL, "bad face_velocity_gradients arguments");
  if (dtype == 0)
    return gradients_impl<double>(c, ny, nx, nvar, static_cast<const double*>(W), dx, dy, axis,
                                  static_cast<double*>(dv_dx), static_cast<double*>(dv_dy));
  return gradients_impl<float>(c, ny, nx, nvar, static_cast<const float*>(W), dx, dy, axis,
                               static_cast<float*>(dv_dx), static_cast<float*>(dv_dy));
}

int jx_smagorinsky_viscosity(jx_ctx* c, int dtype, long long n, const void* dv_dx,
                             const void* dv_dy, double dx, double dy, double smagorinsky_cs,
                             void* nu) {
  if (!c) return JX_EINVAL;
  if (n < 0 || !dv_dx || !dv_dy || !nu) return fail(c, JX_EINVAL, "bad smagorinsky arguments");
  if (n == 0) return JX_OK;
  // (Cs**2 * dx) * dy in float64, then times the strain norm (fluxes.py:245).
  const double k = smagorinsky_cs * smagorinsky_cs * dx * dy;
  const dim3 g = aux_grid((size_t)n);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::smagorinsky_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const double*>(dv_dx), static_cast<const double*>(dv_dy),
                       static_cast<double*>(nu), (size_t)n, k);
  else
    hipLaunchKernelGGL(jx::smagorinsky_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<const float*>(dv_dx), static_cast<const float*>(dv_dy),
                       static_cast<float*>(nu), (size_t)n, (float)k);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_add_momentum_flux(jx_ctx* c, int dtype, long long ncells, int nvar, void* flux,
                         const void* add) {
  if (!c) return JX_EINVAL;
  if (ncells < 0 || nvar < 3 || !flux || !add)
    return fail(c, JX_EINVAL, "bad add_momentum_flux arguments");
  const dim3 g = aux_grid((size_t)ncells);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::add_momentum_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<double*>(flux), static_cast<const double*>(add),
                       (size_t)ncells, nvar);
  else
    hipLaunchKernelGGL(jx::add_momentum_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<float*>(flux), static_cast<const float*>(add),
                       (size_t)ncells, nvar);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_apply_drag(jx_ctx* c, int dtype, long long ncells, int nvar, void* U, const void* W,
                  double beta, double dt, double drag_rate) {
  if (!c) return JX_EINVAL;
  if (ncells < 0 || nvar < 3 || !U || !W) return fail(c, JX_EINVAL, "bad drag arguments");
  const dim3 g = aux_grid((size_t)ncells);
  if (dtype == 0) {
    const double coef = (beta * dt) * drag_rate;
    hipLaunchKernelGGL(jx::drag_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<double*>(U), static_cast<const double*>(W), (size_t)ncells,
                       nvar, coef);
  } else {
    const float coef = (float)(beta * dt) * (float)drag_rate;
    hipLaunchKernelGGL(jx::drag_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<float*>(U), static_cast<const float*>(W), (size_t)ncells,
                       nvar, coef);
  }
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

int jx_apply_kick(jx_ctx* c, int dtype, long long ncells, int nvar, void* U,
                  const void* acceleration, double dt) {
  if (!c) return JX_EINVAL;
  if (ncells < 0 || nvar < 3 || !U || !acceleration)
    return fail(c, JX_EINVAL, "bad kick arguments");
  const dim3 g = aux_grid((size_t)ncells);
  if (dtype == 0)
    hipLaunchKernelGGL(jx::kick_add_kernel<double>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<double*>(U), static_cast<const double*>(acceleration),
                       (size_t)ncells, nvar, dt);
  else
    hipLaunchKernelGGL(jx::kick_add_kernel<float>, g, dim3(jx::kAuxBlock), 0, c->stream,
                       static_cast<float*>(U), static_cast<const float*>(acceleration),
                       (size_t)ncells, nvar, (float)dt);
  JX_HIP(c, hipGetLastError());
  return JX_OK;
}

// ---- batched ensemble engine ----------------------------------------------------------

int jx_batch_create(int device, int dtype, int max_instances, long long pool_bytes,
                    jx_batch** out) {
  if (!out) return JX_EINVAL;
  *out = nullptr;
  if ((dtype != 0 && dtype != 1) || max_instances < 1 || pool_bytes < 256)
    return JX_EINVAL;
  std::string err;
  jx::BatchBase* impl = dtype == 1
      ? jx_make_batch_f32(device, max_instances, pool_bytes, &err)
      : jx_make_batch_f64(device, max_instances, pool_bytes, &err);
  if (!impl) {
    std::fprintf(stderr, "jx_batch_create: %s\n", err.c_str());
    return JX_EHIP;
  }
  *out = new jx_batch{impl, ""};
  return JX_OK;
}

void jx_batch_destroy(jx_batch* b) {
  if (!b) return;
  delete b->impl;
  delete b;
}

const char* jx_batch_last_error(const jx_batch* b) { return b ? b->impl->err.c_str() : "null"; }

long long jx_batch_footprint(const jx_batch* b, const jx_run_params* p, int forced) {
  return (b && p) ? b->impl->footprint(p, forced) : -1;
}

long long jx_batch_pool_free(const jx_batch* b) { return b ? b->impl->pool_free() : -1; }

int jx_batch_add(jx_batch* b, const jx_run_params* p, const void* U0, const jx_forcing* f,
                 int rows, int* id_out) {
  return b ? b->impl->add(p, U0, f, rows, id_out) : JX_EINVAL;
}

int jx_batch_add_resume(jx_batch* b, const jx_run_params* p, const void* U0,
                        const jx_forcing* f, int rows, const jx_loop_state* start, int* id_out) {
  return b ? b->impl->add(p, U0, f, rows, id_out, start) : JX_EINVAL;
}

int jx_batch_checkpoint(jx_batch* b, int id, void* U_out, jx_loop_state* state_out) {
  return b ? b->impl->checkpoint(id, U_out, state_out) : JX_EINVAL;
}

int jx_batch_remove(jx_batch* b, int id) { return b ? b->impl->remove(id) : JX_EINVAL; }

int jx_batch_run(jx_batch* b, int rounds, int* active_out) {
  return b ? b->impl->run(rounds, active_out) : JX_EINVAL;
}

int jx_batch_launch(jx_batch* b, int rounds) {
  return b ? b->impl->launch(rounds) : JX_EINVAL;
}

int jx_batch_wait(jx_batch* b, int* active_out) {
  return b ? b->impl->wait(active_out) : JX_EINVAL;
}

int jx_batch_poll(jx_batch* b, int* done, double* time, int* cycle, int cap) {
  return b ? b->impl->poll(done, time, cycle, cap) : JX_EINVAL;
}

int jx_batch_poll_launched(jx_batch* b, int* done, double* time, int* cycle, int cap,
                           int* active_out) {
  return b ? b->impl->poll_launched(done, time, cycle, cap, active_out) : JX_EINVAL;
}

int jx_batch_result(jx_batch* b, int id, void* U_out, double* time, double* dt, int* cycle) {
  return b ? b->impl->result(id, U_out, time, dt, cycle) : JX_EINVAL;
}

int jx_batch_state(jx_batch* b, int id, void** state_out) {
  return b ? b->impl->state_ptr(id, state_out) : JX_EINVAL;
}

int jx_batch_set_graph(jx_batch* b, int use_graph, int rounds_per_graph) {
  return b ? b->impl->set_graph(use_graph, rounds_per_graph) : JX_EINVAL;
}

int jx_batch_set_fill(jx_batch* b, int fill) {
  return b ? b->impl->set_fill(fill) : JX_EINVAL;
}

}  // extern "C"

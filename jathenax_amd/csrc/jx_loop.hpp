// Device-side loop control shared by the single-grid and batched kernels: the time-loop
// carry of jathenax/run.py:36-68 kept in the Ctrl block and the dt choice of
// hydro.py:68-89.
#pragma once
#include <hip/hip_runtime.h>

#include "jx_ops.hpp"

namespace jx {

__host__ __device__ inline double value_of_max_key(unsigned long long k) {
  if (k == ~0ull) return __builtin_nan("");
  union { unsigned long long u; double d; } c;
  c.u = k;
  return c.d;
}

// NaN-propagating minimum (jnp.minimum).
__host__ __device__ inline double nan_min(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  return a < b ? a : b;
}

// dt from the reduced limits (hydro.py:68-89), with previous dt, time and tlim.
__host__ __device__ inline double finish_timestep(double cfl_limit, double nu_max, int smag_on,
                                                  double viscosity, double drag_rate, double h,
                                                  double cfl, double previous_dt, double time,
                                                  double tlim) {
  double lim = cfl_limit;
  if (smag_on) {
    const double rate = (viscosity + 2.0 * nu_max) / (h * h);
    lim = nan_min(lim, 0.1875 / rate);
  } else if (viscosity > 0.0) {
    lim = nan_min(lim, 0.25 * (h * h) / viscosity);
  }
  if (drag_rate > 0.0) lim = nan_min(lim, 1.0 / drag_rate);
  double dt = nan_min(2.0 * previous_dt, cfl * lim);
  if ((time < tlim) && (time + dt > tlim)) dt = tlim - time;
  return dt;
}

struct LoopArgs {
  double tlim, cfl, viscosity, drag_rate, h, dt_update, tdriv_start, dedt;
  int nlim, smag_on, has_forcing, n_updates;
};

__device__ inline void init_ctrl(Ctrl* c) {
  c->time = 0.0;
  c->dt = 0.0;
  c->prev_dt = __builtin_inf();
  c->cycle = 0;
  c->done = 0;
  c->pending = 0;
  c->update = -1;
  c->render = 0;
  c->power = 0.0;
  c->dtkey = ~0ull;
  c->nukey = 0ull;
  c->mu[0] = c->mu[1] = 0.0;
  c->scale = 0.0;
  c->sum_rho = 0.0;
}

// One thread: commit the previous step, test the loop condition (run.py:36-38), pick dt
// (run.py:42-46) and the OU row due at the cycle-start time (forcing.py:125-134).
__device__ inline void begin_step(Ctrl* c, const LoopArgs& p) {
  if (c->done) return;
  if (c->pending) {
    c->time = c->time + c->dt;
    c->prev_dt = c->dt;
    c->cycle += 1;
    c->pending = 0;
  }
  if (!((c->time < p.tlim) && (c->cycle < p.nlim))) {
    c->done = 1;
    return;
  }
  const double dt = finish_timestep(value_of_key(c->dtkey), value_of_max_key(c->nukey),
                                    p.smag_on, p.viscosity, p.drag_rate, p.h, p.cfl, c->prev_dt,
                                    c->time, p.tlim);
  c->dt = dt;
  c->pending = 1;
  c->dtkey = ~0ull;
  c->nukey = 0ull;
  c->render = 0;
  if (p.has_forcing) {
    int req = -1;
    if (c->time >= p.tdriv_start) req = (int)floor((c->time - p.tdriv_start) / p.dt_update);
    if (req > c->update) {
      c->update = req;
      c->render = 1;
    }
    c->power = (c->update >= 0) ? p.dedt : 0.0;
  }
}

}  // namespace jx

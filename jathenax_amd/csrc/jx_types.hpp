// Plain-old-data types shared by the host launchers and the gfx950 kernels.
#pragma once
#include <cstdint>

namespace jx {

// Reconstruction and Riemann-solver ids, in the order of the reference registries
// (jathenax/fluxes.py:120-123 RECONSTRUCTIONS, fluxes.py:204 RIEMANN_SOLVERS).
enum Recon : int { RECON_DC = 0, RECON_PLM = 1, RECON_PPM4 = 2, RECON_PPMX = 3, RECON_WENOZ = 4 };
enum Rsolver : int { RS_LLF = 0, RS_HLLE = 1, RS_ROE = 2 };

// Device-resident control block of one simulation: the while_loop carry of
// jathenax/run.py:63-68 (time, previous dt, cycle, OU update index) plus the reduction
// slots that feed the next timestep (hydro.py:63-89) and the kick (forcing.py:137-149).
struct Ctrl {
  double time;       // simulation time (float64 in both precisions)
  double dt;         // dt of the step in flight
  double prev_dt;    // dt of the last completed step (+inf before the first)
  int cycle;
  int done;          // set once time >= tlim or cycle >= nlim; every kernel then exits
  int pending;       // a step was launched whose time/cycle are not yet committed
  int update;        // OU row currently rendered (-1 = none)
  int render;        // the raw force must be re-rendered this step
  int pad0;
  double power;      // dedt once the OU index is >= 0, else 0 (run.py:53)
  // CFL limit min over cells, encoded: 0 = NaN, else bits(limit) + 1 (order-preserving).
  unsigned long long dtkey;
  // Smagorinsky nu max over faces, encoded like dtkey (0 = NaN).
  unsigned long long nukey;
  double mu[2];      // mass-weighted mean raw force (forcing.py:141)
  double scale;      // kick scale (forcing.py:146-148)
  double sum_rho;    // box sums used by the kick
  double pad1[2];
};

// Runtime parameters of one isothermal run (the static arguments of
// evolve_isothermal, run.py:24-27, plus grid geometry).
struct RunParams {
  int nx, ny, nvar;
  int recon, rsolver;
  int integrator;            // number of RK stages: 2, 3 or 4
  double dx, dy, cs, dfloor, cfl;
  double viscosity, smag_cs, drag_rate;
  double tlim;
  int nlim;
  int has_forcing;
  double dt_update, tdriv_start, dedt;
  int n_updates, n_modes;
};

// One RK stage's coefficient row (hydro.py:20-36) resolved on the host.
struct StageCoef {
  double gam0, gam1, beta, delta;
  int has_delta;   // RK4 stages 2-4 accumulate the saved register
  int blended;     // gam0 != 0 and gam0 + gam1 == 1.0 in float64 (hydro.py:46)
};

}  // namespace jx

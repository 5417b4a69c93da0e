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

// One RK stage's coefficient row (hydro.py:20-36) resolved on the host.
struct StageCoef {
  double gam0, gam1, beta, delta;
  int has_delta;   // RK4 stages 2-4 accumulate the saved register
  int blended;     // gam0 != 0 and gam0 + gam1 == 1.0 in float64 (hydro.py:46)
};


// Built-in RK rows (hydro.py:20-36) of rk2, rk3 and rk4 by stage count - 2:
// (gam0, gam1, beta, delta) and whether the row has a delta (None otherwise).
struct RkRow { double g0, g1, beta, delta; int has_delta; };
constexpr RkRow kBuiltinRk[3][4] = {
    {{0.0, 1.0, 1.0, 0.0, 0}, {0.5, 0.5, 0.5, 0.0, 0}, {}, {}},
    {{0.0, 1.0, 1.0, 0.0, 0}, {0.25, 0.75, 0.25, 0.0, 0}, {2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 0.0, 0},
     {}},
    {{0.0, 1.0, 1.193743905974738, 0.0, 0},
     {0.121098479554482, 0.721781678111411, 0.099279895495783, 0.217683334308543, 1},
     {-3.843833699660025, 2.121209265338722, 1.131678018054042, 1.065841341361089, 1},
     {0.546370891121863, 0.198653035682705, 0.310665766509336, 0.0, 1}}};

// A caller's table (rows of gam0, gam1, beta, delta with delta NaN for None) equals the
// built-in rows of n stages.
inline bool rk_rows_are_builtin(const double* rows, int n) {
  if (n < 2 || n > 4) return false;
  for (int k = 0; k < n; ++k) {
    const RkRow& b = kBuiltinRk[n - 2][k];
    const double* r = rows + 4 * k;
    const bool none = r[3] != r[3];
    if (r[0] != b.g0 || r[1] != b.g1 || r[2] != b.beta || none == (b.has_delta != 0) ||
        (!none && r[3] != b.delta))
      return false;
  }
  return true;
}

}  // namespace jx

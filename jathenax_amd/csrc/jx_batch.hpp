// Batched ensemble engine: many independent periodic grids of uneven sizes and mixed
// operators advanced together by persistent wave-worker kernels.
//
// Work unit.  A tile is a 64-column strip (58 output columns plus the 3+3 halo) of one
// instance over a band of rows.  Each 64-lane wave is an independent worker: it takes the
// next tile of the round's work list, runs the instance's current RK stage on it with the
// same column march as the single-grid kernel (x exchanges through the wave's own LDS
// slice, so no block barrier couples workers), and moves on.  Reconstruction, Riemann
// solver and viscosity are per-instance runtime choices, uniform across a wave; the
// number of variables selects one of two compiled tile bodies.
//
// Tiles are not stored: tile q of an instance is strip q % nstrips of row band
// q / nstrips.  The band height is chosen every round by the control kernel: the largest
// multiple of 8 rows (at most the instance's cap) that still gives every resident wave
// about `fill` tiles, so a round with little work is cut into small tiles that keep the
// whole GPU busy, and a round with much work uses tall ones that amortise the 6-row
// warm-up of the column march.  Sums over an instance (the kick moments) are taken per
// 8-row group of a strip and then over groups in a fixed order, so they do not depend on
// the band height, and an instance's result does not depend on the rest of the batch.
//
// Rounds.  One round advances every active instance by one RK stage of its own schedule:
// control (advance the stage, begin-step for instances at stage 0, band heights, work
// lists) -> stage (with the forcing render of stage 0) -> kick (mean, moments, scale,
// apply; forced instances at their final stage) -> Smagorinsky timestep.  All control
// lives in device memory, so a round is a fixed sequence of launches that is captured
// once in a hipGraph and replayed.
#pragma once
#include <hip/hip_runtime.h>

#include "jx_aux.hpp"

namespace jx {

constexpr int kBatchNV = 4;            // variables per instance: hydro + one scalar
constexpr int kBatchBlock = 256;       // four wave workers per block
constexpr int kTileCols = 64 - 6;      // output columns per strip
constexpr int kBatchSlice = 4 * kBatchNV + 9;   // LDS rows of 64 values per wave worker

template <typename T> struct BatchInst {
  T* A;                  // state between steps (floored after every step)
  T* B; T* C; T* S;      // stage registers
  T* raw;                // rendered raw force (forced instances)
  double* part;          // kick sums, 3 doubles per (strip, 8-row group), strip-major
  const T* coef; const T* cx; const T* sx; const T* cy; const T* sy;
  const int* grp;        // FP32 render: mode groups of the basis (nullptr: per-mode render)
  int nx, ny, nv, recon, rs, nstages;
  int visc_on, smag_on, drag_on, forced, n_updates, n_modes;
  long long reg_stride;  // elements from A to B, B to C, C to S
  int nstrips, ngroups;  // 58-column strips, 8-row groups
  int rows_cap;          // largest band height (multiple of 8)
  int pad_;
  T dx, dy, cs, cs2, inv_cs, dfloor, eps, four_dx, four_dy, visc, smag_c, drag;
  T inv_dx, inv_dy, inv_four_dx, inv_four_dy;
  LoopArgs loop;
};

struct BatchCtrl {
  Ctrl c;
  int stage;    // RK stage this round executes
  int active;   // takes part in this round
  int live;     // slot holds an instance
  int rows;     // band height of this round's tiles
  int ntiles;   // tiles of this round
  int pad[3];
};

struct Tile { int x0, y0, rows, strip; };

constexpr int kGroupRows = 8;

enum { CNT_STAGE = 0, CNT_MOM, CNT_APPLY, CNT_TSTEP, CNT_TOTAL };

// Per-round work lists, built by the control kernel: the instances that need a kernel
// this round and the prefix sums of their tile counts, so worker waves only visit the
// tiles that have work.  Layout of list L at lists + L * list_stride(cap):
//   [0] entries n, [1] total tiles, [2 .. 2+n] prefix (n+1 values), then n instance ids.
// WL_KICK: instances ending a forced step; WL_SMAG: Smagorinsky instances ending a step
// (their kick, when forced, runs inside the timestep pass); WL_APPLY: forced instances
// without Smagorinsky, whose kick is a pass of its own.
enum { WL_STAGE = 0, WL_KICK, WL_SMAG, WL_APPLY, WL_TOTAL };

// Kick sums per (strip, 8-row group): FP32 takes rho, rho r, rho |r|^2, m.r and m in the
// final stage's own pass (forcing.py:141-145 expanded about the mean, summed in float64);
// FP64 takes rho and rho r there and the second moments in a pass of their own, in the
// reference's form.
template <typename T> constexpr int batch_kick_sums() { return Fast<T>::value ? 7 : 3; }

__host__ __device__ inline int list_stride(int cap) { return 2 + (cap + 1) + cap; }

template <typename T> struct BatchTabs {
  BatchInst<T>* inst;
  BatchCtrl* ctrl;
  int* meta;              // [0] instance slots in use, [2] active count
  unsigned* counters;     // per-round work counters
  int* lists;             // WL_TOTAL work lists
  int* flags;             // per-slot work bits of this round
  int cap;                // instance slots
  int nwaves;             // resident waves of a worker grid
  int fill;               // tiles per wave the band heights aim for
};

// k-th tile of work list L (k < total): instance slot s and its tile q.
template <typename T>
__device__ __forceinline__ void list_item(const BatchTabs<T>& b, const int* L, int k, int& s,
                                          int& q) {
  const cmem_t<int>* Lc = (const cmem_t<int>*)L;
  const int n = Lc[0];
  const cmem_t<int>* pref = Lc + 2;
  const cmem_t<int>* ids = pref + (b.cap + 1);
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  s = ids[lo];
  q = k - pref[lo];
}

// Tile q of an instance at band height `rows`: strip q % nstrips of band q / nstrips.
template <typename T>
__device__ __forceinline__ Tile tile_of(const BatchInst<T>& I, int rows, int q) {
  Tile t;
  t.strip = q % I.nstrips;
  t.x0 = t.strip * kTileCols;
  t.y0 = (q / I.nstrips) * rows;
  t.rows = min(rows, I.ny - t.y0);
  return t;
}

// Register (0 = A) that holds a forced instance's final-stage output before its kick: a
// forced Smagorinsky instance's final stage writes a register that stage does not read, and
// the timestep pass kicks it into A; other forced instances are kicked in A in place.
//   rk2: B|A -> C   rk3: C|A -> B   rk4: B|S -> C
template <typename T> __device__ __forceinline__ int kick_reg(const BatchInst<T>& I) {
  return (I.forced && I.smag_on) ? (I.nstages == 3 ? 1 : 2) : 0;
}

// Fixed-order butterfly sum over the wave of three doubles.
__device__ __forceinline__ void wave_sum3(double& a0, double& a1, double& a2) {
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off);
    a1 += __shfl_xor(a1, off);
    a2 += __shfl_xor(a2, off);
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Next tile index for this wave, as a wave-uniform scalar.  Every lane issues the atomic
// (lane 0 adds 1, the others 0) and the wave takes lane 0's old value, so the loop header
// holds no lane-divergent branch and every exit of the worker loops is a scalar branch.
__device__ __forceinline__ int wave_next(unsigned* counter) {
  const unsigned inc = ((threadIdx.x & 63) == 0) ? 1u : 0u;
  const unsigned old = atomicAdd(counter, inc);
  return __builtin_amdgcn_readfirstlane((int)old);
}

// Work distribution of the persistent kernels: wave w of the grid takes item w first, then
// items nwaves + (counter value) until the list is exhausted.  A kernel whose list is
// shorter than the grid issues no atomic at all, and a long list costs one atomic per item
// beyond the first wave of items.
__device__ __forceinline__ int wave_first() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
}
__device__ __forceinline__ int wave_then(unsigned* counter) {
  return (int)(gridDim.x * (blockDim.x >> 6)) + wave_next(counter);
}

// RK rows (gam0, gam1, beta, delta) of RK2, RK3 and low-storage RK4 by stage
// (hydro.py:20-36), in constant memory so that a runtime stage index reads them with
// scalar loads.
static __constant__ double kRKRows[3][4][4] = {
    {{0.0, 1.0, 1.0, 0.0}, {0.5, 0.5, 0.5, 0.0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
    {{0.0, 1.0, 1.0, 0.0}, {0.25, 0.75, 0.25, 0.0}, {2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 0.0},
     {0, 0, 0, 0}},
    {{0.0, 1.0, 1.193743905974738, 0.0},
     {0.121098479554482, 0.721781678111411, 0.099279895495783, 0.217683334308543},
     {-3.843833699660025, 2.121209265338722, 1.131678018054042, 1.065841341361089},
     {0.546370891121863, 0.198653035682705, 0.310665766509336, 0.0}}};

__device__ __forceinline__ void rk_row(int ns, int k, double& g0, double& g1, double& beta,
                                       double& delta, int& has_delta) {
  const double* row = kRKRows[ns - 2][k];
  g0 = row[0];
  g1 = row[1];
  beta = row[2];
  delta = row[3];
  has_delta = ns == 4 && k > 0;
}

template <typename T>
__device__ __forceinline__ void riemann_rt(int rs, T rl, T rr, T ul, T ur, T tl, T tr,
                                           const Gas<T>& g, T& f0, T& f1, T& f2) {
  // FP32 Roe and LLF fluxes come doubled, as in the fused stage kernel (flux_x2).
  switch (rs) {
    case RS_LLF:
      riemann_x<RS_LLF, false, flux_x2<RS_LLF, T>()>(rl, rr, ul, ur, tl, tr, g, f0, f1, f2);
      break;
    case RS_HLLE: riemann_x<RS_HLLE, false, false>(rl, rr, ul, ur, tl, tr, g, f0, f1, f2); break;
    default:
      riemann_x<RS_ROE, false, flux_x2<RS_ROE, T>()>(rl, rr, ul, ur, tl, tr, g, f0, f1, f2);
      break;
  }
}

template <typename T>
__device__ __forceinline__ T over_rt(T x, T d, T inv) {
  if constexpr (Fast<T>::value) return x * inv;
  else return x / d;
}

// One RK stage of one instance (NV variables, reconstruction RECON) over one tile, executed
// by one wave.  Same arithmetic as stage_kernel; `lds` is this wave's slice
// (kBatchSlice rows of 64).  The Riemann solver and the physics switches stay runtime
// choices (uniform branches); the reconstruction is compiled in, so that the LDS reads and
// the stencil arithmetic of all variables can be scheduled together.
template <typename T, int NV, int RECON>
__device__ __forceinline__ void batch_stage_tile(const BatchInst<T>& I, BatchCtrl& C,
                                                 const Tile& tl, T* lds) {
  // PPM4 shares each clamped interface between the two cells it bounds (in y through a
  // 4-row register ring, in x through the sI slice); the other methods need the full
  // 5-row stencil.
  constexpr bool SHARED = RECON == RECON_PPM4;
  constexpr int R = SHARED ? 4 : 5;
  const int t = threadIdx.x & 63;
  T* sW = lds;             // [NV][64]
  T* sH = sW + NV * 64;    // [NV][64]
  T* sF = sH + NV * 64;    // [NV][64]
  T* sV = sF + NV * 64;    // [3][3][64]
  T* sI = sV + 9 * 64;     // [NV][64] PPM4 x interfaces
  const int nx = I.nx, ny = I.ny;
  constexpr int nv = NV;
  const int x0 = tl.x0;
  int xc = (x0 + t - 3) % nx;
  if (xc < 0) xc += nx;
  const bool col_out = (t >= 3) && (t < 61) && (x0 + t - 3 < nx);
  const int tm2 = max(t - 2, 0), tm1 = max(t - 1, 0), tp1 = min(t + 1, 63), tp2 = min(t + 2, 63);
  const int y0 = tl.y0;
  const int yend = min(y0 + tl.rows, ny);
  const int nit = (yend - y0) + 6;

  // Register plan of the stage (all saved-register reads are pointwise):
  //   rk2: A->B, B|A->A   rk3: A->B, B|A->C, C|A->A
  //   rk4: A->B, B|A->C (S=A+d2 B), C|S->B (S+=d3 C), B|S->A (S+=0 B)
  const int ns = I.nstages, k = C.stage;
  double g0, g1, beta, delta;
  int has_delta;
  rk_row(ns, k, g0, g1, beta, delta, has_delta);
  const bool blended = (g0 != 0.0 && g0 + g1 == 1.0);
  const bool last = (k == ns - 1);
  // Registers A, B, C, S sit reg_stride elements apart, so the plan is kept as register
  // numbers (integer selects) rather than as a choice between pointers.
  int in_r = 0, out_r = 1, sv_r = 0, so_r = 0;
  bool floor_saved = false, has_sv = false;
  if (k > 0) {
    if (ns == 2) { in_r = 1; out_r = 0; }
    else if (ns == 3) { in_r = k; out_r = (k == 1) ? 2 : 0; }
    else {
      in_r = (k == 2) ? 2 : 1;
      out_r = (k == 1) ? 2 : (k == 2) ? 1 : 0;
      so_r = 3;
    }
    const bool from_a = (ns != 4) || (k == 1);
    sv_r = from_a ? 0 : 3;
    floor_saved = from_a;
    has_sv = true;
    if (k == ns - 1) out_r = kick_reg(I);
  }
  gmem_t<T>* const regs = gmem(I.A);
  const gmem_t<T>* g_in = regs + (size_t)in_r * I.reg_stride;
  gmem_t<T>* g_out = regs + (size_t)out_r * I.reg_stride;
  const gmem_t<T>* g_sv = regs + (size_t)sv_r * I.reg_stride;
  gmem_t<T>* g_sout = regs + (size_t)so_r * I.reg_stride;
  const gmem_t<T>* g_raw = gmem(static_cast<const T*>(I.raw));
  gmem_t<double>* g_part = gmem(I.part);
  const bool kick = last && I.forced;
  const bool want_key = last && !I.forced && !I.smag_on;

  const double dt = C.c.dt;
  const T sdt = T(beta * dt);
  const Gas<T> gas{I.cs, I.cs2, I.inv_cs};
  const Visc<T> vp{I.visc, I.smag_c, I.smag_on != 0};
  const T dfloor = I.dfloor;
  const T drag_coef = sdt * I.drag;
  const bool visc_on = I.visc_on != 0;
  // Doubled fluxes (FP32 Roe and LLF): viscous terms at h = 1, update coefficients halved.
  const bool x2 = Fast<T>::value && I.rs != RS_HLLE;
  const T vh = x2 ? T(1) : T(0.5);
  const T cx = flux_coef<T>(beta * dt, I.inv_dx, x2);
  const T cy = flux_coef<T>(beta * dt, I.inv_dy, x2);
  constexpr bool floors = RECON == RECON_PPMX || RECON == RECON_WENOZ;

  T w[R][NV], u[R][NV], hy[NV], gp[NV], gn[NV], pf[NV], ip[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    hy[v] = gp[v] = gn[v] = pf[v] = ip[v] = T(0);
#pragma unroll
    for (int q = 0; q < R; ++q) w[q][v] = u[q][v] = T(0);
  }
  {
    int yn = (y0 - 3) % ny;
    if (yn < 0) yn += ny;
    const gmem_t<T>* src = g_in + ((size_t)yn * nx + xc) * nv;
#pragma unroll
    for (int v = 0; v < NV; ++v)
      pf[v] = src[v];
  }
  // Kick sums of the current 8-row group: rho and rho r in float64; FP32's four one-pass
  // sums (rho |r|^2, m.r, m) in float within the group (at most 8 terms per lane), which
  // keeps the tile within three waves per SIMD.
  constexpr int KS = batch_kick_sums<T>();
  double acc[3] = {0.0, 0.0, 0.0};
  float acc2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  CflMin<T> cfl(I.dx, I.dy, I.cs);

  auto body = [&](auto pc, int i) {
    constexpr int P = decltype(pc)::value;
    constexpr int S0 = P, S1 = (P + R - 1) % R, S2 = (P + R - 2) % R, S3 = (P + R - 3) % R;
    constexpr int S4 = (P + R - 4) % R;
    const int r = y0 - 3 + i;
    // Unconditional loads, saved register first, then the prefetch of row r+1 (see the
    // single-grid stage kernel): the update waits for the first only.
    T svr[NV];
    {
      // Without a saved register (stage 1) the tile's first input row, which stays
      // cached; during the warm-up likewise.
      const gmem_t<T>* sv = has_sv ? g_sv + ((size_t)(i >= 6 ? r - 3 : y0) * nx + xc) * nv
                                   : g_in + ((size_t)y0 * nx + xc) * nv;
#pragma unroll
      for (int v = 0; v < NV; ++v)
        svr[v] = sv[v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) u[S0][v] = pf[v];
    {
      int yn = r + 1;   // periodic row; only the 3-row halos wrap
      if (yn < 0 || yn >= ny) {
        yn %= ny;
        if (yn < 0) yn += ny;
      }
      const gmem_t<T>* src = g_in + ((size_t)yn * nx + xc) * nv;
#pragma unroll
      for (int v = 0; v < NV; ++v)
        pf[v] = src[v];
    }
    {
      const T rho = floor_to(u[S0][0], dfloor);
      u[S0][0] = rho;
      const T inv = dvd(T(1), rho);
      w[S0][0] = rho;
#pragma unroll
      for (int v = 1; v < NV; ++v) w[S0][v] = inv * u[S0][v];
    }
    if (visc_on && i >= 2) {
      T* slot = sV + ((r - 2 + 3 * ny + 3) % 3) * 3 * 64;
      slot[t] = w[S2][0];
      slot[64 + t] = w[S2][1];
      slot[128 + t] = w[S2][2];
    }
    if (i >= 6) {
#pragma unroll
      for (int v = 0; v < NV; ++v) sW[v * 64 + t] = w[S3][v];
    }
    wave_sync();
    // x stencil of row r-3, read before the y work (unused words during the warm-up).
    T xm2[NV], xm1[NV], xp1[NV], xp2[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const T* row = sW + v * 64;
      xm2[v] = row[tm2];
      xm1[v] = row[tm1];
      xp1[v] = row[tp1];
      if constexpr (!SHARED) xp2[v] = row[tp2];
    }

    // y direction: edges of row r-2, flux at its low face.
    auto y_face = [&](T (&lo)[NV], T (&hi)[NV]) {
      if (i >= 5) {
        T f0, f1, f2;
        riemann_rt<T>(I.rs, hy[0], lo[0], hy[2], lo[2], hy[1], lo[1], gas, f0, f1, f2);
        gn[0] = f0;
        gn[1] = f2;
        gn[2] = f1;
#pragma unroll
        for (int v = 3; v < NV; ++v) gn[v] = f0 * ((f0 >= T(0)) ? hy[v] : lo[v]);
        if (visc_on) {
          const T* rs = sV + ((r - 2 + 3 * ny + 3) % 3) * 3 * 64;
          const T* rm = sV + ((r - 3 + 3 * ny + 3) % 3) * 3 * 64;
          const T uy = over_rt(w[S2][1] - w[S3][1], I.dy, I.inv_dy);
          const T vy = over_rt(w[S2][2] - w[S3][2], I.dy, I.inv_dy);
          const T ux = over_rt((rs[64 + tp1] + rm[64 + tp1]) - (rs[64 + tm1] + rm[64 + tm1]),
                               I.four_dx, I.inv_four_dx);
          const T vx = over_rt((rs[128 + tp1] + rm[128 + tp1]) - (rs[128 + tm1] + rm[128 + tm1]),
                               I.four_dx, I.inv_four_dx);
          T fmx_, fmy_;
          visc_face_h<0>(ux, vy, uy, vx, w[S2][0] + w[S3][0], vp, vh, fmx_, fmy_);
          gn[1] = gn[1] + fmx_;
          gn[2] = gn[2] + fmy_;
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) hy[v] = hi[v];
    };
    if constexpr (SHARED) {
      if (i >= 3) {
        T in[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
          in[v] = ppm4_interface(w[S3][v], w[S2][v], w[S1][v], w[S0][v]);   // below row r-1
        if (i >= 4) {
          T lo[NV], hi[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            lo[v] = ip[v];
            hi[v] = in[v];
            ppm4_limit(w[S2][v], lo[v], hi[v]);
          }
          y_face(lo, hi);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) ip[v] = in[v];
      }
    } else {
      if (i >= 4) {
        T lo[NV], hi[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
          cell_edges<RECON>(w[S4][v], w[S3][v], w[S2][v], w[S1][v], w[S0][v], I.eps, lo[v],
                            hi[v]);
        if (floors) {
          lo[0] = fmx(lo[0], dfloor);
          hi[0] = fmx(hi[0], dfloor);
        }
        y_face(lo, hi);
      }
    }

    if (i >= 6) {
      T xl[NV], xh[NV], l[NV];
      if constexpr (SHARED) {
        // Left states from the left neighbour's limiter, as in the single-grid kernel.
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          xl[v] = ppm4_interface(xm2[v], xm1[v], w[S3][v], xp1[v]);
          sI[v * 64 + t] = xl[v];
        }
        wave_sync();
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          T nlo = sI[v * 64 + tm1], nhi = xl[v];
          ppm4_limit(xm1[v], nlo, nhi);
          l[v] = nhi;
          xh[v] = sI[v * 64 + tp1];
          ppm4_limit(w[S3][v], xl[v], xh[v]);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          cell_edges<RECON>(xm2[v], xm1[v], w[S3][v], xp1[v], xp2[v], I.eps, xl[v], xh[v]);
        if (floors) {
          xl[0] = fmx(xl[0], dfloor);
          xh[0] = fmx(xh[0], dfloor);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sH[v * 64 + t] = xh[v];
        wave_sync();
#pragma unroll
        for (int v = 0; v < NV; ++v) l[v] = sH[v * 64 + tm1];
      }
      T fx[NV];
      {
        T f0, f1, f2;
        riemann_rt<T>(I.rs, l[0], xl[0], l[1], xl[1], l[2], xl[2], gas, f0, f1, f2);
        fx[0] = f0;
        fx[1] = f1;
        fx[2] = f2;
#pragma unroll
        for (int v = 3; v < NV; ++v) fx[v] = f0 * ((f0 >= T(0)) ? l[v] : xl[v]);
        if (visc_on) {
          const T* rj = sV + ((r - 3 + 3 * ny + 3) % 3) * 3 * 64;
          const T* rp = sV + ((r - 2 + 3 * ny + 3) % 3) * 3 * 64;
          const T* rq = sV + ((r - 4 + 3 * ny + 3) % 3) * 3 * 64;
          const T ux = over_rt(rj[64 + t] - rj[64 + tm1], I.dx, I.inv_dx);
          const T vx = over_rt(rj[128 + t] - rj[128 + tm1], I.dx, I.inv_dx);
          const T uy = over_rt((rp[64 + t] + rp[64 + tm1]) - (rq[64 + t] + rq[64 + tm1]),
                               I.four_dy, I.inv_four_dy);
          const T vy = over_rt((rp[128 + t] + rp[128 + tm1]) - (rq[128 + t] + rq[128 + tm1]),
                               I.four_dy, I.inv_four_dy);
          T fmx_, fmy_;
          visc_face_h<1>(ux, vy, uy, vx, rj[t] + rj[tm1], vp, vh, fmx_, fmy_);
          fx[1] = fx[1] + fmx_;
          fx[2] = fx[2] + fmy_;
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) sF[v * 64 + t] = fx[v];
      wave_sync();
      if (col_out) {
        const int j = r - 3;
        const size_t cell = (size_t)j * nx + (x0 + t - 3);
        T un[NV], us[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) us[v] = (k == 0) ? u[S3][v] : svr[v];
        if (k > 0 && floor_saved) us[0] = floor_to(us[0], dfloor);
        if (has_delta) {
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            us[v] = us[v] + T(delta) * u[S3][v];
            g_sout[cell * nv + v] = us[v];
          }
        }
        if constexpr (Fast<T>::value) {
          T dF[NV], dG[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            dF[v] = sF[v * 64 + tp1] - fx[v];
            dG[v] = gn[v] - gp[v];
          }
          rk_update_fast<NV>(u[S3], us, dF, dG, k == 0, blended, T(g0), T(g1), cx, cy, un);
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            const T div = over_rt(sF[v * 64 + tp1] - fx[v], I.dx, I.inv_dx) +
                          over_rt(gn[v] - gp[v], I.dy, I.inv_dy);
            if (blended)
              un[v] = us[v] + (T(g0) * (u[S3][v] - us[v]) - sdt * div);
            else
              un[v] = (T(g0) * u[S3][v] + T(g1) * us[v]) - sdt * div;
          }
        }
        if (I.drag_on) {
          __asm__ volatile("");   // a branch, not selects (as in the fused stage kernel)
          const T c = drag_coef * w[S3][0];
          un[1] = un[1] + (-(c * w[S3][1]));
          un[2] = un[2] + (-(c * w[S3][2]));
        }
        if (kick) {
          const gmem_t<T>* rf = g_raw + cell * 2;
          if constexpr (KS == 7) {
            // float products are exact in float64
            const double rho = un[0], r0 = rf[0], r1 = rf[1];
            acc[0] += rho;
            acc[1] += rho * r0;
            acc[2] += rho * r1;
            const float f0 = rf[0], f1 = rf[1];
            acc2[0] += un[0] * (f0 * f0 + f1 * f1);
            acc2[1] += un[1] * f0 + un[2] * f1;
            acc2[2] += un[1];
            acc2[3] += un[2];
          } else {
            acc[0] += (double)un[0];
            acc[1] += (double)(un[0] * rf[0]);
            acc[2] += (double)(un[0] * rf[1]);
          }
        } else {
          un[0] = floor_to(un[0], dfloor);
          if (want_key) {
            const T inv = dvd(T(1), un[0]);
            cfl.add(inv * un[1], inv * un[2]);
          }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
          g_out[cell * nv + v] = un[v];
      }
      // Kick sums close at the end of every 8-row group (bands start on group bounds).
      const int j = r - 3;
      if (kick && (((j & (kGroupRows - 1)) == kGroupRows - 1) || j == yend - 1)) {
        double all[KS];
#pragma unroll
        for (int q = 0; q < KS; ++q) all[q] = (q < 3) ? acc[q] : (double)acc2[q < 3 ? 0 : q - 3];
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
          for (int q = 0; q < KS; ++q) all[q] += __shfl_xor(all[q], off);
        }
        if (t == 0) {
          gmem_t<double>* g = g_part + KS * ((size_t)tl.strip * I.ngroups + j / kGroupRows);
#pragma unroll
          for (int q = 0; q < KS; ++q) g[q] = all[q];
        }
        acc[0] = acc[1] = acc[2] = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc2[q] = 0.0f;
      }
    }
    // gn is set from iteration 5 on and gp first read at iteration 6.
#pragma unroll
    for (int v = 0; v < NV; ++v) gp[v] = gn[v];
  };

  for (int i0 = 0; i0 < nit; i0 += R) {
    body(IC<0>{}, i0);
    if (i0 + 1 < nit) body(IC<1>{}, i0 + 1);
    if (i0 + 2 < nit) body(IC<2>{}, i0 + 2);
    if (i0 + 3 < nit) body(IC<3>{}, i0 + 3);
    if constexpr (R == 5) {
      if (i0 + 4 < nit) body(IC<4 % R>{}, i0 + 4);
    }
  }
  if (want_key) {
    unsigned long long kmin = cfl.key();
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(kmin, off);
      kmin = o < kmin ? o : kmin;
    }
    if (t == 0) atomicMin(&C.c.dtkey, kmin);
  }
}

template <typename T>
__device__ __forceinline__ void batch_render_tile(const BatchInst<T>& I, const BatchCtrl& C,
                                                  const Tile& tl, T* lds);

template <typename T, int NV>
__device__ __forceinline__ void batch_stage_dispatch(const BatchInst<T>& I, BatchCtrl& C,
                                                     const Tile& tl, T* lds) {
  switch (I.recon) {
    case RECON_DC: batch_stage_tile<T, NV, RECON_DC>(I, C, tl, lds); break;
    case RECON_PLM: batch_stage_tile<T, NV, RECON_PLM>(I, C, tl, lds); break;
    case RECON_PPM4: batch_stage_tile<T, NV, RECON_PPM4>(I, C, tl, lds); break;
    case RECON_PPMX: batch_stage_tile<T, NV, RECON_PPMX>(I, C, tl, lds); break;
    default: batch_stage_tile<T, NV, RECON_WENOZ>(I, C, tl, lds); break;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBatchBlock, 2) batch_stage_kernel(BatchTabs<T> b) {
  static_assert(kBatchSlice * 64 >= 4 * kWaveGroupChunk * 8, "render slice too small");
  extern __shared__ double bsm[];
  T* lds = reinterpret_cast<T*>(bsm) + (threadIdx.x >> 6) * kBatchSlice * 64;
  const int* L = b.lists + WL_STAGE * list_stride(b.cap);
  const int total = ((const cmem_t<int>*)L)[1];
  for (int k = wave_first(); k < total; k = wave_then(&b.counters[CNT_STAGE])) {
    int s, q;
    list_item(b, L, k, s, q);
    // By-value copy: the tile's stores to the state could alias the instance record, so
    // a reference would make the compiler reload its fields on every row.
    const BatchInst<T> I = load_const(b.inst + s);
    BatchCtrl& C = b.ctrl[s];
    const Tile tl = tile_of(I, C.rows, q);
    if (I.nv == 4) batch_stage_dispatch<T, 4>(I, C, tl, lds);
    else batch_stage_dispatch<T, 3>(I, C, tl, lds);
    // A step that opened a new forcing row renders the raw force of this tile (read by
    // the kick at the step's final stage).
    if (C.stage == 0 && I.forced && C.c.render) batch_render_tile<T>(I, C, tl, lds);
  }
}

// Control, one block of kCtrlBlock threads: advance the stage of last round's workers,
// open steps (begin-step) for instances at stage 0, pick this round's band height, and
// build the work lists (slot order, block-wide scans in LDS) and reset the counters.
constexpr int kCtrlBlock = 256;

template <typename T>
__global__ void __launch_bounds__(kCtrlBlock) batch_control_kernel(BatchTabs<T> b) {
  __shared__ int sc[WL_TOTAL][2][kCtrlBlock];
  __shared__ int base[WL_TOTAL][2];
  __shared__ long long red[kCtrlBlock];
  const int tid = threadIdx.x;
  if (tid < CNT_TOTAL) b.counters[tid] = 0u;
  if (tid < WL_TOTAL) base[tid][0] = base[tid][1] = 0;
  const int n = b.meta[0];
  long long work = 0;   // strip-rows of this round
  for (int s = tid; s < n; s += kCtrlBlock) {
    BatchCtrl& C = b.ctrl[s];
    const BatchInst<T>& I = b.inst[s];
    int f = 0;
    if (C.live) {
      if (C.active) C.stage = (C.stage + 1) % I.nstages;
      if (C.stage == 0) begin_step(&C.c, I.loop);
      C.active = !C.c.done;
      if (C.active) {
        const bool last = C.stage == I.nstages - 1;
        f = (1 << WL_STAGE) | ((last && I.forced) ? 1 << WL_KICK : 0) |
            ((last && I.smag_on) ? 1 << WL_SMAG : 0) |
            ((last && I.forced && !I.smag_on) ? 1 << WL_APPLY : 0);
        work += (long long)I.nstrips * I.ny;
      }
    } else {
      C.active = 0;
    }
    b.flags[s] = f;
  }
  red[tid] = work;
  __syncthreads();
  for (int off = kCtrlBlock / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const long long target = (long long)b.nwaves * b.fill;
  long long band = target > 0 ? red[0] / target : (1 << 20);   // fill 0: every cap
  band = band < (1 << 20) ? band : (1 << 20);
  const int rows_round = max(kGroupRows, (int)band / kGroupRows * kGroupRows);

  for (int c0 = 0; c0 < n; c0 += kCtrlBlock) {
    const int s = c0 + tid;
    int f = 0, nt = 0;
    if (s < n) {
      f = b.flags[s];
      if (f) {
        BatchCtrl& C = b.ctrl[s];
        const BatchInst<T>& I = b.inst[s];
        const int rows = min(I.rows_cap, rows_round);
        nt = I.nstrips * ((I.ny + rows - 1) / rows);
        C.rows = rows;
        C.ntiles = nt;
      }
    }
#pragma unroll
    for (int l = 0; l < WL_TOTAL; ++l) {
      const int on = (f >> l) & 1;
      sc[l][0][tid] = on;
      sc[l][1][tid] = on ? nt : 0;
    }
    __syncthreads();
    for (int off = 1; off < kCtrlBlock; off <<= 1) {   // inclusive Hillis-Steele scans
      int v[WL_TOTAL][2];
#pragma unroll
      for (int l = 0; l < WL_TOTAL; ++l)
#pragma unroll
        for (int j = 0; j < 2; ++j) v[l][j] = tid >= off ? sc[l][j][tid - off] : 0;
      __syncthreads();
#pragma unroll
      for (int l = 0; l < WL_TOTAL; ++l)
#pragma unroll
        for (int j = 0; j < 2; ++j) sc[l][j][tid] += v[l][j];
      __syncthreads();
    }
#pragma unroll
    for (int l = 0; l < WL_TOTAL; ++l) {
      if ((f >> l) & 1) {
        int* L = b.lists + l * list_stride(b.cap);
        const int k = base[l][0] + sc[l][0][tid] - 1;
        L[2 + k] = base[l][1] + sc[l][1][tid] - nt;          // prefix
        L[2 + (b.cap + 1) + k] = s;                          // instance id
      }
    }
    __syncthreads();
    if (tid < WL_TOTAL) {
      base[tid][0] += sc[tid][0][kCtrlBlock - 1];
      base[tid][1] += sc[tid][1][kCtrlBlock - 1];
    }
    __syncthreads();
  }
  if (tid < WL_TOTAL) {
    int* L = b.lists + tid * list_stride(b.cap);
    L[0] = base[tid][0];
    L[1] = base[tid][1];
    L[2 + base[tid][0]] = base[tid][1];
    if (tid == WL_STAGE) b.meta[2] = base[tid][0];
  }
}

template <typename T>
__device__ __forceinline__ void batch_render_tile(const BatchInst<T>& I, const BatchCtrl& C,
                                                  const Tile& tl, T* lds) {
  const int t = threadIdx.x & 63;
  const int row = C.c.update;
  const bool fill = (row < 0) || (row >= I.n_updates);
  const T* amp = I.coef + (size_t)(fill ? 0 : row) * I.n_modes * 4;
  const int i = tl.x0 + t;
  const bool on = t < kTileCols && i < I.nx;   // per lane; the row loop stays uniform
  if constexpr (Fast<T>::value) {
    if (I.grp) {
      for (int yb = tl.y0; yb < tl.y0 + tl.rows; yb += 8)
        render_wave_grouped(gmem(I.raw), amp, fill, I.cx, I.sx, I.cy, I.sy, I.grp, I.nx, I.ny,
                            I.n_modes, on ? i : 0, on, yb, min(8, tl.y0 + tl.rows - yb), lds);
      return;
    }
  }
  for (int yb = tl.y0; yb < tl.y0 + tl.rows; yb += 8)
    render_wave<T>(gmem(I.raw), amp, fill, I.cx, I.sx, I.cy, I.sy, I.nx, I.ny, I.n_modes,
                   on ? i : 0, on, yb, min(8, tl.y0 + tl.rows - yb), lds);
}

__device__ __forceinline__ bool kicking(const BatchCtrl& C, int forced, int nstages) {
  return C.active && forced && C.stage == nstages - 1;
}

// Per-instance fixed-order sums of the group partials, one wave per instance slot: lane l
// adds groups l, l + 64, ... in order, then a fixed butterfly.  mode 0: mean force (FP32:
// and the scale, from the one-pass sums); mode 1 (FP64): scale from the second moments.
template <typename T>
__global__ void __launch_bounds__(256) batch_kick_reduce_kernel(BatchTabs<T> b, int mode) {
  const int s = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  if (s >= b.meta[0]) return;
  BatchCtrl& C = b.ctrl[s];
  const BatchInst<T>& I = b.inst[s];
  if (!kicking(C, I.forced, I.nstages)) return;
  const int ng = I.nstrips * I.ngroups;
  const gmem_t<double>* part = gmem(static_cast<const double*>(I.part));
  // FP64 second moments (mode 1) come in threes from the moment pass.
  const int K = (mode == 0) ? batch_kick_sums<T>() : 3;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int g = lane; g < ng; g += 64) {
#pragma unroll
    for (int q = 0; q < 7; ++q)
      if (q < K) a[q] += part[(size_t)K * g + q];
  }
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 7; ++q) a[q] += __shfl_xor(a[q], off);
  }
  if (lane != 0) return;
  const double ncell = (double)I.nx * I.ny;
  double m0, m1;
  if (mode == 0) {
    C.c.sum_rho = a[0];
    C.c.mu[0] = a[1] / a[0];
    C.c.mu[1] = a[2] / a[0];
    if (K != 7) return;
    // sum rho |r - mu|^2 = sum rho |r|^2 - mu . sum rho r; sum m.(r - mu) likewise
    const double mu0 = C.c.mu[0], mu1 = C.c.mu[1];
    const double var = a[3] - (mu0 * a[1] + mu1 * a[2]);
    m0 = 0.5 * C.c.dt * var / ncell;
    m1 = (a[4] - (mu0 * a[5] + mu1 * a[6])) / ncell;
  } else {
    m0 = a[0] / ncell;
    m1 = a[1] / ncell;
  }
  {
    const double dedt = C.c.power;
    const double rt = sqrt(m1 * m1 + 4.0 * m0 * dedt);
    double scale = (m1 >= 0.0) ? 2.0 * dedt / (m1 + rt) : (rt - m1) / (2.0 * m0);
    if (!(m0 > 1e-30)) scale = __builtin_nan("");
    C.c.scale = scale;
  }
}

// Second moments (mode 0, FP64) or the kick itself (mode 1, forced instances without
// Smagorinsky) on one tile of a kicking instance.  The moments read the final stage's output
// wherever it is (kick_reg); the kick works on A in place.
template <typename T>
__device__ void batch_kick_tile(const BatchInst<T>& I, BatchCtrl& C, const Tile& tl, int mode) {
  const int t = threadIdx.x & 63;
  const T mu0 = T(C.c.mu[0]), mu1 = T(C.c.mu[1]);
  const T dt = T(C.c.dt);
  const bool off = C.c.power == 0.0;
  const T scale = T(C.c.scale);
  const int i = tl.x0 + t;
  const bool on = t < kTileCols && i < I.nx;
  const int yend = tl.y0 + tl.rows;
  double s0 = 0, s1 = 0, s2 = 0;
  CflMin<T> cfl(I.dx, I.dy, I.cs);
  gmem_t<T>* A = gmem(I.A) + (mode == 0 ? (size_t)kick_reg(I) * I.reg_stride : 0);
  const gmem_t<T>* raw = gmem(static_cast<const T*>(I.raw));
  gmem_t<double>* part = gmem(I.part);
  // Rows are walked by every lane (uniform loop); lanes past the strip or the grid skip.
  for (int j = tl.y0; j < yend; ++j) {
    const size_t c = (size_t)j * I.nx + i;
    gmem_t<T>* q = A + c * I.nv;
    if (mode == 0) {
      if (on) {
        const T f0 = raw[c * 2] - mu0, f1 = raw[c * 2 + 1] - mu1;
        const T fsq = f0 * f0 + f1 * f1;
        s0 += (double)(((T(0.5) * q[0]) * fsq) * dt);
        s1 += (double)(q[1] * f0 + q[2] * f1);
      }
      if (((j & (kGroupRows - 1)) == kGroupRows - 1) || j == yend - 1) {
        wave_sum3(s0, s1, s2);
        if (t == 0) {
          gmem_t<double>* g = part + 3 * ((size_t)tl.strip * I.ngroups + j / kGroupRows);
          g[0] = s0;
          g[1] = s1;
          g[2] = 0.0;
        }
        s0 = s1 = s2 = 0.0;
      }
    } else if (on) {
      const T f0 = raw[c * 2] - mu0, f1 = raw[c * 2 + 1] - mu1;
      const T acc0 = off ? T(0) : scale * f0;
      const T acc1 = off ? T(0) : scale * f1;
      const T rho = q[0];
      const T m0 = q[1] + rho * (acc0 * dt);
      const T m1 = q[2] + rho * (acc1 * dt);
      const T rf = floor_to(rho, I.dfloor);
      q[0] = rf;
      q[1] = m0;
      q[2] = m1;
      if (!I.smag_on) {
        const T inv = dvd(T(1), rf);
        cfl.add(inv * m0, inv * m1);
      }
    }
  }
  if (mode == 1 && !I.smag_on) {
    unsigned long long kmin = cfl.key();
    for (int off2 = 32; off2 > 0; off2 >>= 1) {
      const unsigned long long o = __shfl_xor(kmin, off2);
      kmin = o < kmin ? o : kmin;
    }
    if (t == 0) atomicMin(&C.c.dtkey, kmin);
  }
}

template <typename T>
__global__ void __launch_bounds__(kBatchBlock) batch_kick_tiles_kernel(BatchTabs<T> b, int mode) {
  const int* L = b.lists + (mode == 0 ? WL_KICK : WL_APPLY) * list_stride(b.cap);
  const int total = ((const cmem_t<int>*)L)[1];
  unsigned* ctr = &b.counters[mode == 0 ? CNT_MOM : CNT_APPLY];
  for (int k = wave_first(); k < total; k = wave_then(ctr)) {
    int s, q;
    list_item(b, L, k, s, q);
    const BatchInst<T> I = load_const(b.inst + s);   // by value (see batch_stage_kernel)
    BatchCtrl& C = b.ctrl[s];
    batch_kick_tile<T>(I, C, tile_of(I, C.rows, q), mode);
  }
}

// CFL and Smagorinsky limits over one tile of an instance's current state A. Lane t follows
// column x0 - 1 + t (periodic) and marches up the tile's rows; lanes 1..kTileCols are the
// tile's columns, lanes 0 and kTileCols + 1 their neighbours. A cell's primitives are
// recovered once, as in the single-grid timestep kernel, and the face expressions are the
// same, so both give the same limits bit for bit.
//
// `kick` (a forced Smagorinsky instance ending a step): the rows come from the final stage's
// output (kick_reg) and each row is kicked as it enters, with batch_kick_tile's expressions
// (mu, scale and dt from the control block); the lanes that own a cell of the tile's rows
// write its kicked state to A, and the limits are those of the kicked state.  Halo rows and
// columns are kicked by every tile that reads them, with the owner's arithmetic.
template <typename T>
__device__ void batch_timestep_tile(const BatchInst<T>& I, BatchCtrl& C, const Tile& tl,
                                    bool kick) {
  const int t = threadIdx.x & 63;
  const int nx = I.nx, ny = I.ny, nv = I.nv;
  int xc = (tl.x0 - 1 + t) % nx;
  if (xc < 0) xc += nx;
  const bool on = t >= 1 && t <= kTileCols && tl.x0 + t - 1 < nx;
  CflMin<T> cfl(I.dx, I.dy, I.cs);
  unsigned long long kmax = 0ull;
  gmem_t<T>* const A = gmem(I.A);
  const gmem_t<T>* src = A + (kick ? (size_t)kick_reg(I) * I.reg_stride : 0);
  const gmem_t<T>* raw = gmem(static_cast<const T*>(I.raw));
  const int j0 = tl.y0, j1 = min(tl.y0 + tl.rows, ny);
  T mu0 = T(0), mu1 = T(0), kdt = T(0), scale = T(0);
  bool off = true;
  if (kick) {
    mu0 = T(C.c.mu[0]);
    mu1 = T(C.c.mu[1]);
    kdt = T(C.c.dt);
    off = C.c.power == 0.0;
    scale = T(C.c.scale);
  }
  // A row's raw values at this lane's column: the state, and with the kick the raw force and
  // the tracer (nvar 4) that the kicked copy carries.
  struct Row {
    T q[3];
    T f[2];
    T x;
  };
  auto fetch = [&](int j, Row& w) {
    if (j < 0 || j >= ny) {
      j %= ny;
      if (j < 0) j += ny;
    }
    const size_t cell = (size_t)j * nx + xc;
    const gmem_t<T>* q = src + cell * nv;
    w.q[0] = q[0];
    w.q[1] = q[1];
    w.q[2] = q[2];
    w.f[0] = w.f[1] = w.x = T(0);
    if (kick) {
      w.f[0] = raw[cell * 2];
      w.f[1] = raw[cell * 2 + 1];
      if (nv > 3) w.x = q[3];
    }
  };
  // Row j enters: kick it (the owned cells are written to A), then recover.
  auto prim = [&](int j, Row& w, T& u, T& v) {
    if (kick) {
      const T f0 = w.f[0] - mu0, f1 = w.f[1] - mu1;
      const T acc0 = off ? T(0) : scale * f0;
      const T acc1 = off ? T(0) : scale * f1;
      const T rho = w.q[0];
      const T m0 = w.q[1] + rho * (acc0 * kdt);
      const T m1 = w.q[2] + rho * (acc1 * kdt);
      w.q[0] = floor_to(rho, I.dfloor);
      w.q[1] = m0;
      w.q[2] = m1;
      if (on && j >= j0 && j < j1) {
        gmem_t<T>* d = A + ((size_t)j * nx + xc) * nv;
        d[0] = w.q[0];
        d[1] = m0;
        d[2] = m1;
        if (nv > 3) d[3] = w.x;
      }
    }
    T r;
    load_prim(w.q, nv, false, I.dfloor, r, u, v);
  };
  // Rows are fetched two ahead of their use (see the single-grid kernel).
  Row q1, q2;
  fetch(j0, q1);
  fetch(j0 + 1, q2);
  if (!I.smag_on) {
    for (int j = j0; j < j1; ++j) {
      T u, v;
      prim(j, q1, u, v);
      q1 = q2;
      fetch(j + 2, q2);
      if (on) cfl.add(u, v);
    }
  } else {
    T ou_s, ov_s, ou_c, ov_c;
    {
      Row q0;
      fetch(j0 - 1, q0);
      prim(j0 - 1, q0, ou_s, ov_s);
    }
    prim(j0, q1, ou_c, ov_c);
    q1 = q2;
    fetch(j0 + 2, q2);
    T lu_s = __shfl_up(ou_s, 1), lv_s = __shfl_up(ov_s, 1);
    T lu_c = __shfl_up(ou_c, 1), lv_c = __shfl_up(ov_c, 1);
    T ru_s = __shfl_down(ou_s, 1), rv_s = __shfl_down(ov_s, 1);
    T ru_c = __shfl_down(ou_c, 1), rv_c = __shfl_down(ov_c, 1);
    for (int j = j0; j < j1; ++j) {
      T ou_n, ov_n;
      prim(j + 1, q1, ou_n, ov_n);
      q1 = q2;
      fetch(j + 3, q2);
      const T lu_n = __shfl_up(ou_n, 1), lv_n = __shfl_up(ov_n, 1);
      const T ru_n = __shfl_down(ou_n, 1), rv_n = __shfl_down(ov_n, 1);
      if (on) {
        const T vx = ou_c, vy = ov_c;
        cfl.add(vx, vy);
        const T u_w = lu_c, v_w = lv_c, u_n = ou_n, v_n = ov_n, u_nw = lu_n, v_nw = lv_n;
        const T u_s = ou_s, v_s = ov_s, u_sw = lu_s, v_sw = lv_s;
        const T u_e = ru_c, v_e = rv_c, u_se = ru_s, v_se = rv_s;
        {
          const T ux = dvd(vx - u_w, I.dx);
          const T vxn = dvd(vy - v_w, I.dx);
          const T uy = dvd((u_n + u_nw) - (u_s + u_sw), I.four_dy);
          const T vyt = dvd((v_n + v_nw) - (v_s + v_sw), I.four_dy);
          const unsigned long long k = max_key_of((double)smag_nu(ux, vyt, vxn + uy, I.smag_c));
          kmax = k > kmax ? k : kmax;
        }
        {
          const T uy = dvd(vx - u_s, I.dy);
          const T vyn = dvd(vy - v_s, I.dy);
          const T uxt = dvd((u_e + u_se) - (u_w + u_sw), I.four_dx);
          const T vxt = dvd((v_e + v_se) - (v_w + v_sw), I.four_dx);
          const unsigned long long k = max_key_of((double)smag_nu(uxt, vyn, vxt + uy, I.smag_c));
          kmax = k > kmax ? k : kmax;
        }
      }
      ou_s = ou_c; ov_s = ov_c; ou_c = ou_n; ov_c = ov_n;
      lu_s = lu_c; lv_s = lv_c; lu_c = lu_n; lv_c = lv_n;
      ru_s = ru_c; rv_s = rv_c; ru_c = ru_n; rv_c = rv_n;
    }
  }
  unsigned long long kmin = cfl.key();
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(kmin, off);
    kmin = o < kmin ? o : kmin;
    const unsigned long long p = __shfl_xor(kmax, off);
    kmax = p > kmax ? p : kmax;
  }
  if (t == 0) {   // skip updates that cannot change the slot (see timestep_kernel)
    if (kmin < *reinterpret_cast<volatile unsigned long long*>(&C.c.dtkey))
      atomicMin(&C.c.dtkey, kmin);
    if (I.smag_on && kmax > *reinterpret_cast<volatile unsigned long long*>(&C.c.nukey))
      atomicMax(&C.c.nukey, kmax);
  }
}

// `only` >= 0 selects one instance (its initial timestep); otherwise Smagorinsky instances
// that have just finished a step.
template <typename T>
__global__ void __launch_bounds__(kBatchBlock) batch_timestep_kernel(BatchTabs<T> b, int only) {
  const int* L = b.lists + WL_SMAG * list_stride(b.cap);
  int total = ((const cmem_t<int>*)L)[1];
  if (only >= 0) {
    const BatchInst<T>& I = b.inst[only];
    total = I.nstrips * ((I.ny + I.rows_cap - 1) / I.rows_cap);
  }
  for (int k = wave_first(); k < total; k = wave_then(&b.counters[CNT_TSTEP])) {
    int s = only, q = k;
    if (only < 0) list_item(b, L, k, s, q);
    const BatchInst<T> I = load_const(b.inst + s);   // by value (see batch_stage_kernel)
    BatchCtrl& C = b.ctrl[s];
    // Within a round (only < 0) the forced Smagorinsky instances are kicked here.
    batch_timestep_tile<T>(I, C, tile_of(I, only >= 0 ? I.rows_cap : C.rows, q),
                           only < 0 && I.forced && I.smag_on);
  }
}

template <typename T>
__global__ void batch_init_kernel(BatchTabs<T> b, int slot) {
  BatchCtrl& C = b.ctrl[slot];
  init_ctrl(&C.c);
  C.stage = 0;
  C.active = 0;
  C.live = 1;
  b.counters[CNT_TSTEP] = 0u;
}

// A resumed instance: the loop state it was checkpointed in.
template <typename T>
__global__ void batch_resume_kernel(BatchTabs<T> b, int slot, double time, double prev_dt,
                                    int cycle, int update) {
  Ctrl& c = b.ctrl[slot].c;
  c.time = time;
  c.prev_dt = prev_dt;
  c.cycle = cycle;
  c.update = update;
}

// Raw force of one instance's current OU row over all its tiles (a resumed forced
// instance); waves take tiles by id, no counters.
template <typename T>
__global__ void __launch_bounds__(kBatchBlock) batch_render_slot_kernel(BatchTabs<T> b,
                                                                        int slot) {
  static_assert(2 * kRenderModes * 8 >= 4 * kWaveGroupChunk * 8, "render slice too small");
  __shared__ __align__(16) T slices[kBatchBlock / 64][2 * kRenderModes * 8];
  const BatchInst<T> I = load_const(b.inst + slot);
  const BatchCtrl& C = b.ctrl[slot];
  const int rows = I.rows_cap;
  const int total = I.nstrips * ((I.ny + rows - 1) / rows);
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  T* lds = slices[threadIdx.x >> 6];
  for (int k = wave_first(); k < total; k += nw)
    batch_render_tile<T>(I, C, tile_of(I, rows, k), lds);
}

template <typename T>
__global__ void batch_retire_kernel(BatchTabs<T> b, int slot) {
  b.ctrl[slot].live = 0;
  b.ctrl[slot].active = 0;
}

static __global__ void reset_counter_kernel(unsigned* c) { *c = 0u; }

}  // namespace jx

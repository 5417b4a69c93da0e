// Auxiliary gfx950 kernels around the fused stage: timestep reduction, the device-side
// loop control, forcing render / normalisation / kick, and the pointwise operators the
// reference exports on their own.
#pragma once
#include <hip/hip_runtime.h>

#include "jx_ops.hpp"
#include "jx_stage.hpp"

namespace jx {

constexpr int kAuxBlock = 256;

// ---------------------------------------------------------------------------------------
// Timestep reduction (hydro.py:63-80): min over cells of dx/(|vx|+cs), dy/(|vy|+cs) and,
// with Smagorinsky, max over all faces of nu_SGS.  `primitive` = the input already holds
// W (isothermal_timestep's argument); otherwise it holds U and is recovered on read.
// ---------------------------------------------------------------------------------------

template <typename T, typename Ptr>   // Ptr: const T* in any address space
__device__ __forceinline__ void load_prim(Ptr q, int nv, bool primitive, T dfloor, T& r,
                                          T& vx, T& vy) {
  (void)nv;
  if (primitive) {
    r = q[0];
    vx = q[1];
    vy = q[2];
  } else {
    r = floor_to(q[0], dfloor);
    const T inv = dvd(T(1), r);
    vx = inv * q[1];
    vy = inv * q[2];
  }
}

template <typename T> struct TimestepArgs {
  const T* q;
  int nx, ny, nv, primitive, smag_on;
  T dx, dy, cs, dfloor, four_dx, four_dy, smag_c;
  const Ctrl* ctrl;               // nullptr: always run
  unsigned long long* dtkey;      // min slot
  unsigned long long* nukey;      // max slot
  // timestep_kernel<T, true> only: q holds the post-stage state of a forced step, raw its
  // raw force (or force.pq its per-row group sums), and the kicked state goes to out
  // (another buffer than q).
  const T* raw;
  T* out;
  int waves;                      // target wave count of the tiling (ts_tiling)
  ForceEval<T> force;
};

__device__ __forceinline__ unsigned long long max_key_of(double x) {
  if (x != x) return ~0ull;
  return static_cast<unsigned long long>(__double_as_longlong(x));
}

// Work decomposition of timestep_kernel: each 64-lane wave owns a strip of kTsCols output
// columns (lanes 1..62; lanes 0 and 63 carry the neighbouring columns) and a band of rows,
// with about 8192 waves in all (32 per CU) so that other waves cover each row's load.
constexpr int kTsCols = 62;
constexpr int kTsAhead = 4;   // rows in flight per lane in the Smagorinsky march
constexpr int kTsWaves = 8192;       // waves of a timestep reduction (32 per CU)
constexpr int kTsKickWaves = 4096;   // with the kick: one resident round at 4 waves per SIMD
struct TsTiling {
  int nstrips, nbands, rows;
};
__host__ __device__ inline TsTiling ts_tiling(int nx, int ny, int waves = 8192) {
  const int nstrips = (nx + kTsCols - 1) / kTsCols;
  int nbands = (waves + nstrips - 1) / nstrips;
  nbands = nbands > ny ? ny : (nbands < 1 ? 1 : nbands);
  const int rows = (ny + nbands - 1) / nbands;
  return {nstrips, (ny + rows - 1) / rows, rows};
}
inline dim3 ts_grid(int nx, int ny, int waves = 8192) {
  const TsTiling t = ts_tiling(nx, ny, waves);
  const int used = t.nstrips * t.nbands;
  return dim3((unsigned)((used + kAuxBlock / 64 - 1) / (kAuxBlock / 64)));
}

// Each lane marches up its column. A cell's primitives are recovered once, when its row
// enters; the row above and below stay in registers and the neighbouring columns come from
// the adjacent lanes (one shuffle per value and row), so the Smagorinsky face stencils cost
// no further loads. The face expressions are those of fluxes.py:227-245 in the reference
// order, so the limits match the cell-by-cell form bit for bit.
//
// KICK: the kick that ends a forced step (forcing.py:149, hydro.py:127-132; mu, scale and
// dt from the control block) is applied to each row as it enters, before its primitives are
// recovered, and the lanes that own a cell write its kicked state to `out`.  Halo rows and
// columns are kicked by every wave that reads them, with the owner's arithmetic, so the
// limits are those of the kicked state; `out` must differ from `q`, which neighbouring waves
// are still reading.  The expressions are kick_apply_kernel's, so the state matches the
// separate kick bit for bit.
template <typename T, bool KICK = false>
__global__ void __launch_bounds__(kAuxBlock) timestep_kernel(TimestepArgs<T> a) {
  if (a.ctrl && a.ctrl->done) return;
  const int nx = a.nx, ny = a.ny, nv = a.nv;
  const TsTiling tl = ts_tiling(nx, ny, a.waves);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)(blockIdx.x * (kAuxBlock / 64) + (threadIdx.x >> 6)));
  CflMin<T> cfl(a.dx, a.dy, a.cs);
  unsigned long long kmax = 0ull;
  T mu0 = T(0), mu1 = T(0), kdt = T(0), scale = T(0);
  bool off = true;
  if constexpr (KICK) {
    mu0 = T(a.ctrl->mu[0]);
    mu1 = T(a.ctrl->mu[1]);
    kdt = T(a.ctrl->dt);
    off = a.ctrl->power == 0.0;
    scale = T(a.ctrl->scale);
  }
  if (wave < tl.nstrips * tl.nbands) {
    const int strip = wave % tl.nstrips, band = wave / tl.nstrips;
    const int x0 = strip * kTsCols;   // first output column
    int xc = (x0 - 1 + lane) % nx;
    if (xc < 0) xc += nx;
    const bool out = lane >= 1 && lane <= kTsCols && x0 + lane - 1 < nx;
    const int j0 = band * tl.rows, j1 = min(j0 + tl.rows, ny);
    ForceCols<T> fc{};
    if constexpr (KICK) {
      if (a.force.pq) fc = force_cols(a.force, nx, xc);
    }
    // Raw values of a row (periodic): the state, and with KICK the raw force and the
    // passive variables (nvar <= 5) that the kicked copy carries.
    struct Row {
      T q[3];
      T f[KICK ? 2 : 1];
      T x[KICK ? 2 : 1];
    };
    auto fetch = [&](int j, Row& w) {
      if (j < 0 || j >= ny) {
        j %= ny;
        if (j < 0) j += ny;
      }
      const size_t cell = (size_t)j * nx + xc;
      const T* src = a.q + cell * nv;
      if constexpr (KICK) {
        // one vector load per lane and array (rows as buffer resources)
        const int row_bytes = nx * nv * (int)sizeof(T);
        const unsigned lane_off = (unsigned)(xc * nv * (int)sizeof(T));
        if (nv == 3) {
          row_load<T, 3>(row_rsrc(a.q + (size_t)j * nx * 3, row_bytes), lane_off, w.q);
          w.x[0] = w.x[1] = T(0);
        } else {
          w.q[0] = src[0];
          w.q[1] = src[1];
          w.q[2] = src[2];
          w.x[0] = src[3];
          w.x[1] = nv > 4 ? src[4] : T(0);
        }
        if (a.force.pq)
          force_at(a.force, fc, j, w.f[0], w.f[1]);
        else
          row_load<T, 2>(row_rsrc(a.raw + (size_t)j * nx * 2, nx * 2 * (int)sizeof(T)),
                         (unsigned)(xc * 2 * (int)sizeof(T)), w.f);
      } else {
        w.q[0] = src[0];
        w.q[1] = src[1];
        w.q[2] = src[2];
      }
    };
    // Row j enters: kick it (KICK; the owned cells are written), then recover.
    auto prim = [&](int j, Row& w, T& u, T& v) {
      if constexpr (KICK) {
        const T f0 = w.f[0] - mu0, f1 = w.f[1] - mu1;
        const T acc0 = off ? T(0) : scale * f0;
        const T acc1 = off ? T(0) : scale * f1;
        const T rho = w.q[0];
        const T m0 = w.q[1] + rho * (acc0 * kdt);
        const T m1 = w.q[2] + rho * (acc1 * kdt);
        w.q[0] = floor_to(rho, a.dfloor);
        w.q[1] = m0;
        w.q[2] = m1;
        if (out && j >= j0 && j < j1) {
          if (nv == 3) {
            row_store<T, 3>(row_rsrc(a.out + (size_t)j * nx * 3,
                                                   nx * 3 * (int)sizeof(T)),
                                          (unsigned)(xc * 3 * (int)sizeof(T)), w.q);
          } else {
            T* dst = a.out + ((size_t)j * nx + xc) * nv;
            dst[0] = w.q[0];
            dst[1] = m0;
            dst[2] = m1;
            dst[3] = w.x[0];
            if (nv > 4) dst[4] = w.x[1];
          }
        }
      }
      T r;
      load_prim(w.q, nv, a.primitive != 0, a.dfloor, r, u, v);
    };
    // Rows are fetched two ahead of their use, so their latency overlaps the work of the
    // rows in between.
    Row q1, q2;
    fetch(j0, q1);
    fetch(j0 + 1, q2);
    if (!a.smag_on) {
      for (int j = j0; j < j1; ++j) {
        T u, v;
        prim(j, q1, u, v);
        q1 = q2;
        fetch(j + 2, q2);
        if (out) cfl.add(u, v);
      }
    } else {
      // Rows s = j-1, c = j of this column (o), the one to the left (l) and to the right
      // (r): velocities vx (u) and vy (v).  Rows j0 .. j0+kTsAhead-1 are fetched up front
      // into a ring (row r in slot (r - j0) mod kTsAhead); each row's slot is refilled with
      // the row kTsAhead further up as soon as the row has entered.  The loop runs whole
      // rounds of kTsAhead rows, so every slot index is a constant and no register copy has
      // to wait for a load in flight.
      constexpr int D = kTsAhead;
      Row ring[D];
      ring[0] = q1;
      ring[1] = q2;
#pragma unroll
      for (int k = 2; k < D; ++k) fetch(j0 + k, ring[k]);
      T ou_s, ov_s, ou_c, ov_c;
      {
        Row q0;
        fetch(j0 - 1, q0);
        prim(j0 - 1, q0, ou_s, ov_s);
      }
      prim(j0, ring[0], ou_c, ov_c);
      fetch(j0 + D, ring[0]);
      T lu_s = __shfl_up(ou_s, 1), lv_s = __shfl_up(ov_s, 1);
      T lu_c = __shfl_up(ou_c, 1), lv_c = __shfl_up(ov_c, 1);
      T ru_s = __shfl_down(ou_s, 1), rv_s = __shfl_down(ov_s, 1);
      T ru_c = __shfl_down(ou_c, 1), rv_c = __shfl_down(ov_c, 1);
      // Row j+1 enters from slot S = (j + 1 - j0) mod D; the faces of row j follow.
      auto row = [&](auto slot, int j) {
        constexpr int S = decltype(slot)::value;
        T ou_n, ov_n;   // row j+1
        prim(j + 1, ring[S], ou_n, ov_n);
        fetch(j + 1 + D, ring[S]);
        const T lu_n = __shfl_up(ou_n, 1), lv_n = __shfl_up(ov_n, 1);
        const T ru_n = __shfl_down(ou_n, 1), rv_n = __shfl_down(ov_n, 1);
        if (out) {
          const T vx = ou_c, vy = ov_c;
          cfl.add(vx, vy);
          const T u_w = lu_c, v_w = lv_c, u_n = ou_n, v_n = ov_n, u_nw = lu_n, v_nw = lv_n;
          const T u_s = ou_s, v_s = ov_s, u_sw = lu_s, v_sw = lv_s;
          const T u_e = ru_c, v_e = rv_c, u_se = ru_s, v_se = rv_s;
          {   // low x face (fluxes.py:227-238, axis=1)
            const T ux = dvd(vx - u_w, a.dx);
            const T vxn = dvd(vy - v_w, a.dx);
            const T uy = dvd((u_n + u_nw) - (u_s + u_sw), a.four_dy);
            const T vyt = dvd((v_n + v_nw) - (v_s + v_sw), a.four_dy);
            const unsigned long long k = max_key_of((double)smag_nu(ux, vyt, vxn + uy, a.smag_c));
            kmax = k > kmax ? k : kmax;
          }
          {   // low y face (axis=0)
            const T uy = dvd(vx - u_s, a.dy);
            const T vyn = dvd(vy - v_s, a.dy);
            const T uxt = dvd((u_e + u_se) - (u_w + u_sw), a.four_dx);
            const T vxt = dvd((v_e + v_se) - (v_w + v_sw), a.four_dx);
            const unsigned long long k = max_key_of((double)smag_nu(uxt, vyn, vxt + uy, a.smag_c));
            kmax = k > kmax ? k : kmax;
          }
        }
        ou_s = ou_c; ov_s = ov_c; ou_c = ou_n; ov_c = ov_n;
        lu_s = lu_c; lv_s = lv_c; lu_c = lu_n; lv_c = lv_n;
        ru_s = ru_c; rv_s = rv_c; ru_c = ru_n; rv_c = rv_n;
      };
      static_assert(D == 4, "the round below is written for a 4-row ring");
      int j = j0;
      for (; j + D <= j1; j += D) {
        row(IC<1>{}, j);
        row(IC<2>{}, j + 1);
        row(IC<3>{}, j + 2);
        row(IC<0>{}, j + 3);
      }
      if (j < j1) row(IC<1>{}, j);
      if (j + 1 < j1) row(IC<2>{}, j + 1);
      if (j + 2 < j1) row(IC<3>{}, j + 2);
    }
  }
  unsigned long long kmin = cfl.key();
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(kmin, off);
    kmin = o < kmin ? o : kmin;
    const unsigned long long q = __shfl_xor(kmax, off);
    kmax = q > kmax ? q : kmax;
  }
  // Waves of a block combine in LDS; then one update per block, skipped when the slot
  // already holds a better value (slots only move one way, so a stale read can only cause
  // a redundant atomic), which keeps thousands of blocks from queueing on two addresses.
  __shared__ unsigned long long sred[2 * (kAuxBlock / 64)];
  const int w = threadIdx.x >> 6;
  if (lane == 0) {
    sred[w] = kmin;
    sred[kAuxBlock / 64 + w] = kmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kAuxBlock / 64; ++k) {
      kmin = sred[k] < kmin ? sred[k] : kmin;
      kmax = sred[kAuxBlock / 64 + k] > kmax ? sred[kAuxBlock / 64 + k] : kmax;
    }
    if (kmin < *reinterpret_cast<volatile unsigned long long*>(a.dtkey)) atomicMin(a.dtkey, kmin);
    if (a.smag_on && kmax > *reinterpret_cast<volatile unsigned long long*>(a.nukey))
      atomicMax(a.nukey, kmax);
  }
}

static __global__ void init_ctrl_kernel(Ctrl* c) { init_ctrl(c); }

static __global__ void begin_step_kernel(Ctrl* c, LoopArgs p) { begin_step(c, p); }

// Resumed run: the loop state another run ended with (time, last dt, cycle, OU row).
static __global__ void set_loop_state_kernel(Ctrl* c, double time, double prev_dt, int cycle,
                                             int update) {
  c->time = time;
  c->prev_dt = prev_dt;
  c->cycle = cycle;
  c->update = update;
}

// Standalone step: dt (and the kick power) come from the caller.
static __global__ void set_step_kernel(Ctrl* c, double dt, double power) {
  c->dt = dt;
  c->done = 0;
  c->power = power;
}

// ---------------------------------------------------------------------------------------
// Forcing: render of the raw acceleration (forcing.py:111-122), normalisation
// (forcing.py:137-149) and the kick (hydro.py:127-132).
// ---------------------------------------------------------------------------------------

template <typename T> struct RenderArgs {
  T* raw;
  const T* table;    // (n_updates, M, 2, 2) OU coefficients in the state dtype
  const T* cx; const T* sx; const T* cy; const T* sy;   // (M, nx), (M, ny)
  int nx, ny, n_modes, n_updates;
  const Ctrl* ctrl;  // nullptr: render row `row` unconditionally
  int row;
};

// The raw force of one wave's 64 columns over rows y0 .. y0+nr-1 (nr <= 8).  Modes go in
// chunks of up to kRenderModes: the wave stages the chunk's row tables (cy, sy of its rows)
// in its LDS slice `lds` (2 * kRenderModes * 8 values), then each lane loads its column's
// cx, sx once per mode and accumulates the 8 rows in registers.  Each cell sums its modes in
// order with the reference's expression (forcing.py:115-122), so the result is the same bit
// for bit as a per-cell loop.  Lanes with on == false follow the uniform control flow.
constexpr int kRenderModes = 32;

__device__ __forceinline__ void render_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, typename PR>
__device__ __forceinline__ void render_wave(PR raw, const T* amp_, bool fill, const T* cx_,
                                            const T* sx_, const T* cy_, const T* sy_, int nx,
                                            int ny, int n_modes, int i, bool on, int y0, int nr,
                                            T* lds) {
  const gmem_t<T>* cx = gmem(cx_);
  const gmem_t<T>* sx = gmem(sx_);
  const gmem_t<T>* cy = gmem(cy_);
  const gmem_t<T>* sy = gmem(sy_);
  const cmem_t<T>* amp = (const cmem_t<T>*)amp_;   // uniform, read-only: scalar loads
  const int lane = threadIdx.x & 63;
  const T nan = T(__builtin_nan(""));
  T f0[8], f1[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) f0[r] = f1[r] = T(0);
  for (int m0 = 0; m0 < n_modes; m0 += kRenderModes) {
    const int mc = min(kRenderModes, n_modes - m0);
    render_wave_sync();   // the previous chunk's reads are done
    for (int e = lane; e < mc * 8; e += 64) {
      const int q = e >> 3, r = e & 7;
      const size_t at = (size_t)(m0 + q) * ny + (y0 + (r < nr ? r : 0));
      lds[2 * e + 0] = cy[at];
      lds[2 * e + 1] = sy[at];
    }
    render_wave_sync();
    for (int q = 0; q < mc; ++q) {
      const int m = m0 + q;
      const T cxv = on ? cx[(size_t)m * nx + i] : T(0);
      const T sxv = on ? sx[(size_t)m * nx + i] : T(0);
      const cmem_t<T>* am = amp + (size_t)m * 4;   // [dir][re/im]
      const T a00 = fill ? nan : am[0], a01 = fill ? nan : am[1];
      const T a10 = fill ? nan : am[2], a11 = fill ? nan : am[3];
      const T* ly = lds + 16 * q;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (r < nr) {
          const T cyv = ly[2 * r], syv = ly[2 * r + 1];
          const T re = cxv * cyv - sxv * syv;
          const T im = cxv * syv + sxv * cyv;
          f0[r] = f0[r] + (re * a00 - im * a01);
          f1[r] = f1[r] + (re * a10 - im * a11);
        }
      }
    }
  }
  if (on) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r < nr) {
        raw[((size_t)(y0 + r) * nx + i) * 2 + 0] = f0[r];
        raw[((size_t)(y0 + r) * nx + i) * 2 + 1] = f1[r];
      }
    }
  }
}

// Rows of one render block; blocks are kAuxBlock columns wide (four waves, each with its own
// LDS slice) and kRenderRows rows tall, or 8 rows when the grid is small.
constexpr int kRenderRows = 16;

__host__ __device__ inline int render_rows_per_block(int nx, int ny) {
  const long long blocks = (long long)((nx + 255) / 256) * ((ny + 15) / 16);
  return blocks >= 1024 ? kRenderRows : 8;
}

template <typename T>
__global__ void __launch_bounds__(kAuxBlock) render_force_kernel(RenderArgs<T> a) {
  __shared__ T slices[kAuxBlock / 64][2 * kRenderModes * 8];
  int row = a.row;
  if (a.ctrl) {
    if (a.ctrl->done || !a.ctrl->render) return;
    row = a.ctrl->update;
  }
  const bool fill = (row < 0) || (row >= a.n_updates);
  const T* amp = a.table + (size_t)(fill ? 0 : row) * a.n_modes * 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < a.nx;
  const int rb = render_rows_per_block(a.nx, a.ny);
  const int y0 = blockIdx.y * rb;
  T* lds = slices[threadIdx.x >> 6];
  for (int yb = y0; yb < min(y0 + rb, a.ny); yb += 8)
    render_wave<T>(gmem(a.raw), amp, fill, a.cx, a.sx, a.cy, a.sy, a.nx, a.ny, a.n_modes,
                   on ? i : 0, on, yb, min(8, a.ny - yb), lds);
}

// ---------------------------------------------------------------------------------------
// FP32 render by x-wavenumber groups.  Each mode term of forcing.py:115-122 is
//   re a_c0 - im a_c1 = cx (cy a_c0 - sy a_c1) - sx (sy a_c0 + cy a_c1),
// so modes whose x basis rows are equal (cx, sx), or equal up to the sign of sx (the -kx
// partner of kx), share one x factor:
//   f_c(i, j) = sum_g cx_g(i) P_gc(j) - sx_g(i) Q_gc(j),
//   P_gc = sum_{m in g} cy_m a_mc0 - sy_m a_mc1,  Q_gc = sum_{m in g} s_m (sy_m a_mc0 + cy_m a_mc1)
// with s_m = +-1 the sign of the mode's sx against its group's.  The per-row sums P, Q cost
// a few operations per row and mode; each cell then costs 4 multiply-adds per group instead
// of 8 per mode (the Cartesian band 1 <= |k| <= 3 has 28 modes in 4 groups).  Groups are
// found from the tables themselves (exact comparison), so any basis is rendered correctly;
// one with no shared rows has one group per mode.  FP64 keeps the per-mode loop above, whose
// sum order is the reference's.
//
// Group table (ints), for M modes: [0] group count G, [1 .. G+1] first member of each group
// in `order` (G+1 entries), then order[M] (modes by group; the first member's tables are the
// group's), sign[M] (+1 / -1 per mode), and scratch gid[M], rep[M], rel[M*M].
// ---------------------------------------------------------------------------------------

// Offset of rel[] in the group table: G, gstart[M+1], order, sign, gid, rep.
__host__ __device__ inline size_t mode_rel_offset(int M) { return 2 + 5 * (size_t)M; }
__host__ __device__ inline size_t mode_group_ints(int M) {
  return mode_rel_offset(M) + (size_t)M * M;
}

// rel[m * M + n] for n < m: 1 when modes m and n have equal x rows, 2 when cx is equal and sx
// negated, 0 otherwise.  One block per (m, n) pair.
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) mode_relation_kernel(const T* cx, const T* sx,
                                                                  int nx, int M, int* grp) {
  const int m = blockIdx.y, n = blockIdx.x;
  if (n >= m) return;
  int same = 1, anti = 1;
  for (int i = threadIdx.x; i < nx; i += blockDim.x) {
    const T c0 = cx[(size_t)m * nx + i], c1 = cx[(size_t)n * nx + i];
    const T s0 = sx[(size_t)m * nx + i], s1 = sx[(size_t)n * nx + i];
    same &= (c0 == c1) && (s0 == s1);
    anti &= (c0 == c1) && (s0 == -s1);
  }
  same = __syncthreads_and(same);
  anti = __syncthreads_and(anti);
  int* rel = grp + mode_rel_offset(M);
  if (threadIdx.x == 0) rel[(size_t)m * M + n] = same ? 1 : (anti ? 2 : 0);
}

// One thread: each mode joins the first earlier group whose first member it matches.
static __global__ void mode_groups_kernel(int M, int* grp) {
  int* gstart = grp + 1;
  int* order = gstart + (M + 1);
  int* sign = order + M;
  int* gid = sign + M;
  int* rep = gid + M;
  const int* rel = grp + mode_rel_offset(M);
  int G = 0;
  for (int m = 0; m < M; ++m) {
    int g = -1, s = 1;
    for (int h = 0; h < G && g < 0; ++h) {
      const int r = rel[(size_t)m * M + rep[h]];
      if (r) {
        g = h;
        s = (r == 1) ? 1 : -1;
      }
    }
    if (g < 0) {
      g = G++;
      rep[g] = m;
    }
    gid[m] = g;
    sign[m] = s;
  }
  for (int g = 0; g <= G; ++g) gstart[g] = 0;
  for (int m = 0; m < M; ++m) ++gstart[gid[m] + 1];
  for (int g = 0; g < G; ++g) gstart[g + 1] += gstart[g];
  for (int g = 0; g < G; ++g) rep[g] = gstart[g];   // fill cursor
  for (int m = 0; m < M; ++m) order[rep[gid[m]]++] = m;
  grp[0] = G;
}

constexpr int kGroupChunk = 32;

// Row j's sums of group g (P0, Q0, P1, Q1): the y factors of the group's modes with their
// OU amplitudes (NaN when `fill`: no OU row in range), each mode's sx sign folded into Q.
__device__ __forceinline__ float4 group_row_sums(const RenderArgs<float>& a, const int* grp,
                                                 const float* amp, bool fill, int g, int j) {
  const int M = a.n_modes, ny = a.ny;
  const int* gstart = grp + 1;
  const int* order = gstart + (M + 1);
  const int* sign = order + M;
  const float nan = __builtin_nanf("");
  float p0 = 0.0f, q0 = 0.0f, p1 = 0.0f, q1 = 0.0f;
  // Members in chunks of 8: the chunk's basis values are loaded first, then accumulated in
  // member order (the order fixes the sums' rounding).
  constexpr int C = 8;
  const int k0 = gstart[g], k1 = gstart[g + 1];
  for (int kc = k0; kc < k1; kc += C) {
    int mm[C];
    float cyv[C], syv[C];
#pragma unroll
    for (int u = 0; u < C; ++u) {
      mm[u] = (kc + u < k1) ? order[kc + u] : order[kc];
      cyv[u] = a.cy[(size_t)mm[u] * ny + j];
      syv[u] = a.sy[(size_t)mm[u] * ny + j];
    }
#pragma unroll
    for (int u = 0; u < C; ++u) {
      if (kc + u < k1) {
        const int m = mm[u];
        const float s = (float)sign[m];
        const float a00 = fill ? nan : amp[m * 4 + 0], a01 = fill ? nan : amp[m * 4 + 1];
        const float a10 = fill ? nan : amp[m * 4 + 2], a11 = fill ? nan : amp[m * 4 + 3];
        p0 = fmaf(cyv[u], a00, fmaf(-syv[u], a01, p0));
        q0 = fmaf(s, fmaf(syv[u], a00, cyv[u] * a01), q0);
        p1 = fmaf(cyv[u], a10, fmaf(-syv[u], a11, p1));
        q1 = fmaf(s, fmaf(syv[u], a10, cyv[u] * a11), q1);
      }
    }
  }
  return make_float4(p0, q0, p1, q1);
}

// Per-row group sums for force_at (at most kMaxForceGroups groups; the rest of a row's
// slots are zero): one thread per (group slot, row), rows fastest so that the basis rows
// cy, sy are read coalesced.  With a control block, only a step that opens a new OU row
// recomputes them, as the render.
static __global__ void __launch_bounds__(kAuxBlock) force_rows_kernel(RenderArgs<float> a,
                                                                      const int* grp,
                                                                      float4* pq) {
  int row = a.row;
  if (a.ctrl) {
    if (a.ctrl->done || !a.ctrl->render) return;
    row = a.ctrl->update;
  }
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ny * kMaxForceGroups) return;
  const int g = e / a.ny, j = e - g * a.ny;
  const bool fill = (row < 0) || (row >= a.n_updates);
  const float* amp = a.table + (size_t)(fill ? 0 : row) * a.n_modes * 4;
  pq[(size_t)j * kMaxForceGroups + g] = (g < grp[0]) ? group_row_sums(a, grp, amp, fill, g, j)
                                                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Blocks as render_force_kernel: kAuxBlock columns by render_rows_per_block rows.
static __global__ void __launch_bounds__(kAuxBlock) render_grouped_kernel(RenderArgs<float> a,
                                                                   const int* grp) {
  __shared__ float4 pq[kRenderRows][kGroupChunk];   // (P_g0, Q_g0, P_g1, Q_g1)
  int row = a.row;
  if (a.ctrl) {
    if (a.ctrl->done || !a.ctrl->render) return;
    row = a.ctrl->update;
  }
  const int M = a.n_modes, nx = a.nx, ny = a.ny;
  const bool fill = (row < 0) || (row >= a.n_updates);
  const float* amp = a.table + (size_t)(fill ? 0 : row) * M * 4;
  const int G = grp[0];
  const int* gstart = grp + 1;
  const int* order = gstart + (M + 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < nx;
  const int rb = render_rows_per_block(nx, ny);
  const int y0 = blockIdx.y * rb;
  const int nr = min(rb, ny - y0);
  float f0[kRenderRows], f1[kRenderRows];
#pragma unroll
  for (int r = 0; r < kRenderRows; ++r) f0[r] = f1[r] = 0.0f;
  for (int g0 = 0; g0 < G; g0 += kGroupChunk) {
    const int gc = min(kGroupChunk, G - g0);
    __syncthreads();   // the previous chunk's sums are consumed
    for (int e = threadIdx.x; e < gc * rb; e += blockDim.x) {
      const int q = e / rb, r = e - q * rb;
      const int j = y0 + min(r, nr - 1);
      pq[r][q] = group_row_sums(a, grp, amp, fill, g0 + q, j);
    }
    __syncthreads();
    for (int q = 0; q < gc; ++q) {
      const int m = order[gstart[g0 + q]];
      const float cxv = on ? a.cx[(size_t)m * nx + i] : 0.0f;
      const float sxv = on ? a.sx[(size_t)m * nx + i] : 0.0f;
#pragma unroll
      for (int r = 0; r < kRenderRows; ++r) {
        if (r < nr) {
          const float4 v = pq[r][q];
          f0[r] = fmaf(cxv, v.x, fmaf(-sxv, v.y, f0[r]));
          f1[r] = fmaf(cxv, v.z, fmaf(-sxv, v.w, f1[r]));
        }
      }
    }
  }
  if (on) {
    gmem_t<float>* raw = gmem(a.raw);
#pragma unroll
    for (int r = 0; r < kRenderRows; ++r) {
      if (r < nr) {
        raw[((size_t)(y0 + r) * nx + i) * 2 + 0] = f0[r];
        raw[((size_t)(y0 + r) * nx + i) * 2 + 1] = f1[r];
      }
    }
  }
}

// The grouped render for one wave's 64 columns over rows y0 .. y0+nr-1 (nr <= 8), as
// render_wave: the per-row sums of up to kWaveGroupChunk groups go to the wave's LDS slice
// `lds` (16-byte aligned, 4 * kWaveGroupChunk * 8 floats), then each lane adds its column.
constexpr int kWaveGroupChunk = 16;

__device__ __forceinline__ void render_wave_grouped(gmem_t<float>* raw, const float* amp,
                                                    bool fill, const float* cx,
                                                    const float* sx, const float* cy,
                                                    const float* sy, const int* grp, int nx,
                                                    int ny, int M, int i, bool on, int y0,
                                                    int nr, float* lds) {
  const int lane = threadIdx.x & 63;
  const float nan = __builtin_nanf("");
  const int G = grp[0];
  const int* gstart = grp + 1;
  const int* order = gstart + (M + 1);
  const int* sign = order + M;
  float f0[8], f1[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) f0[r] = f1[r] = 0.0f;
  float4* pq = reinterpret_cast<float4*>(lds);   // [group][row]
  for (int g0 = 0; g0 < G; g0 += kWaveGroupChunk) {
    const int gc = min(kWaveGroupChunk, G - g0);
    render_wave_sync();   // the previous chunk's sums are consumed
    for (int e = lane; e < gc * 8; e += 64) {
      const int q = e >> 3, r = e & 7;
      const int j = y0 + (r < nr ? r : 0);
      float p0 = 0.0f, q0 = 0.0f, p1 = 0.0f, q1 = 0.0f;
      for (int k = gstart[g0 + q]; k < gstart[g0 + q + 1]; ++k) {
        const int m = order[k];
        const float sg = (float)sign[m];
        const float cyv = cy[(size_t)m * ny + j], syv = sy[(size_t)m * ny + j];
        const float a00 = fill ? nan : amp[m * 4 + 0], a01 = fill ? nan : amp[m * 4 + 1];
        const float a10 = fill ? nan : amp[m * 4 + 2], a11 = fill ? nan : amp[m * 4 + 3];
        p0 = fmaf(cyv, a00, fmaf(-syv, a01, p0));
        q0 = fmaf(sg, fmaf(syv, a00, cyv * a01), q0);
        p1 = fmaf(cyv, a10, fmaf(-syv, a11, p1));
        q1 = fmaf(sg, fmaf(syv, a10, cyv * a11), q1);
      }
      pq[e] = make_float4(p0, q0, p1, q1);
    }
    render_wave_sync();
    for (int q = 0; q < gc; ++q) {
      const int m = order[gstart[g0 + q]];
      const float cxv = on ? cx[(size_t)m * nx + i] : 0.0f;
      const float sxv = on ? sx[(size_t)m * nx + i] : 0.0f;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (r < nr) {
          const float4 v = pq[q * 8 + r];
          f0[r] = fmaf(cxv, v.x, fmaf(-sxv, v.y, f0[r]));
          f1[r] = fmaf(cxv, v.z, fmaf(-sxv, v.w, f1[r]));
        }
      }
    }
  }
  if (on) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r < nr) {
        raw[((size_t)(y0 + r) * nx + i) * 2 + 0] = f0[r];
        raw[((size_t)(y0 + r) * nx + i) * 2 + 1] = f1[r];
      }
    }
  }
}

template <typename T> struct KickArgs {
  T* u;                 // post-drag state (ny, nx, nv), updated in place by the kick
  const T* raw;
  int nx, ny, nv;
  T dfloor, dx, dy, cs;
  const Ctrl* ctrl;
  double* partial;      // per-block partial sums
  unsigned long long* dtkey;   // CFL slot for the next timestep (nullptr: none)
};

// Partial sums of rho, rho*r over a fixed cell range per block (standalone normalise
// entry; the fused path gets these from the final stage kernel).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) kick_first_moments_kernel(KickArgs<T> a) {
  if (a.ctrl && a.ctrl->done) return;
  __shared__ double sred[3 * (kAuxBlock / 64)];
  const size_t n = (size_t)a.nx * a.ny;
  double s0 = 0, s1 = 0, s2 = 0;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T rho = a.u[c * a.nv];
    s0 += (double)rho;
    s1 += (double)(rho * a.raw[c * 2]);
    s2 += (double)(rho * a.raw[c * 2 + 1]);
  }
  block_reduce_store<T>(s0, s1, s2, sred, a.partial + 3 * (size_t)blockIdx.x);
}

// Sum per-block partials in a fixed order.  mode 0: mean force mu = sum(rho r)/sum(rho)
// (forcing.py:141).  mode 1: m0, m1 and the kick scale (forcing.py:144-148).  mode 2 (the
// FP32 single pass, 7 sums per block: rho, rho r, rho |r|^2, m.r, m): mu, then
//   sum rho |r - mu|^2 = sum rho |r|^2 - mu . sum rho r,   sum m.(r - mu) = sum m.r - mu . sum m,
// and the scale from them.
// One block of kReduceBlock threads: each thread sums a strided share of the partials in a
// fixed order, then a fixed tree in LDS.
constexpr int kReduceBlock = 1024;
static __global__ void __launch_bounds__(kReduceBlock)
reduce_kick_kernel(Ctrl* c, const double* partial, int nblocks, int mode, double ncells) {
  if (c->done) return;
  constexpr int K = 7;
  __shared__ double s[K][kReduceBlock];
  const int nk = mode == 2 ? 7 : 3;
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (k < nk) a[k] += partial[(size_t)nk * b + k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) s[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < K; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double m0, m1;
    if (mode == 0) {
      c->sum_rho = s[0][0];
      c->mu[0] = s[1][0] / s[0][0];
      c->mu[1] = s[2][0] / s[0][0];
      return;
    } else if (mode == 1) {
      m0 = s[0][0] / ncells;
      m1 = s[1][0] / ncells;
    } else {
      const double mu0 = s[1][0] / s[0][0], mu1 = s[2][0] / s[0][0];
      c->sum_rho = s[0][0];
      c->mu[0] = mu0;
      c->mu[1] = mu1;
      const double var = s[3][0] - (mu0 * s[1][0] + mu1 * s[2][0]);
      m0 = 0.5 * c->dt * var / ncells;
      m1 = (s[4][0] - (mu0 * s[5][0] + mu1 * s[6][0])) / ncells;
    }
    const double dedt = c->power;
    const double rt = sqrt(m1 * m1 + 4.0 * m0 * dedt);
    double scale = (m1 >= 0.0) ? 2.0 * dedt / (m1 + rt) : (rt - m1) / (2.0 * m0);
    if (!(m0 > 1e-30)) scale = __builtin_nan("");
    c->scale = scale;
  }
}

// Partial sums of 0.5 rho |f|^2 dt and m.f with f = r - mu (forcing.py:142-145).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) kick_second_moments_kernel(KickArgs<T> a) {
  if (a.ctrl->done) return;
  __shared__ double sred[3 * (kAuxBlock / 64)];
  const T mu0 = T(a.ctrl->mu[0]), mu1 = T(a.ctrl->mu[1]);
  const T dt = T(a.ctrl->dt);
  const size_t n = (size_t)a.nx * a.ny;
  double s0 = 0, s1 = 0;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T* q = a.u + c * a.nv;
    const T f0 = a.raw[c * 2] - mu0, f1 = a.raw[c * 2 + 1] - mu1;
    const T fsq = f0 * f0 + f1 * f1;
    s0 += (double)(((T(0.5) * q[0]) * fsq) * dt);
    s1 += (double)(q[1] * f0 + q[2] * f1);
  }
  block_reduce_store<T>(s0, s1, 0.0, sred, a.partial + 3 * (size_t)blockIdx.x);
}

// Apply the kick, floor the density and feed the next timestep's CFL reduction.
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) kick_apply_kernel(KickArgs<T> a) {
  if (a.ctrl->done) return;
  __shared__ unsigned long long sred[kAuxBlock / 64];
  const T mu0 = T(a.ctrl->mu[0]), mu1 = T(a.ctrl->mu[1]);
  const T dt = T(a.ctrl->dt);
  const bool off = a.ctrl->power == 0.0;
  const T scale = T(a.ctrl->scale);
  const size_t n = (size_t)a.nx * a.ny;
  CflMin<T> cfl(a.dx, a.dy, a.cs);
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    T* q = a.u + c * a.nv;
    const T f0 = a.raw[c * 2] - mu0, f1 = a.raw[c * 2 + 1] - mu1;
    const T acc0 = off ? T(0) : scale * f0;
    const T acc1 = off ? T(0) : scale * f1;
    const T rho = q[0];
    const T m0 = q[1] + rho * (acc0 * dt);
    const T m1 = q[2] + rho * (acc1 * dt);
    const T rf = floor_to(rho, a.dfloor);
    q[0] = rf;
    q[1] = m0;
    q[2] = m1;
    if (a.dtkey) {
      const T inv = dvd(T(1), rf);
      cfl.add(inv * m0, inv * m1);
    }
  }
  if (a.dtkey) block_min_key(cfl.key(), sred, a.dtkey);
}

// Standalone normalize_force output: the acceleration field (forcing.py:149).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) kick_acceleration_kernel(const Ctrl* c,
                                                                      const T* raw, T* out,
                                                                      size_t n) {
  const T mu0 = T(c->mu[0]), mu1 = T(c->mu[1]);
  const bool off = c->power == 0.0;
  const T scale = T(c->scale);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    out[2 * k] = off ? T(0) : scale * (raw[2 * k] - mu0);
    out[2 * k + 1] = off ? T(0) : scale * (raw[2 * k + 1] - mu1);
  }
}

// ---------------------------------------------------------------------------------------
// Pointwise operators exported by the reference (hydro.py:39-60, fluxes.py:269-287) and
// the float64 box sums of run.py:186-187.
// ---------------------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kAuxBlock) recover_kernel(const T* u, T* uf, T* w, size_t n,
                                                            int nv, T dfloor) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T* q = u + c * nv;
    const T rho = floor_to(q[0], dfloor);
    const T inv = dvd(T(1), rho);
    if (uf) {
      uf[c * nv] = rho;
      for (int v = 1; v < nv; ++v) uf[c * nv + v] = q[v];
    }
    if (w) {
      w[c * nv] = rho;
      for (int v = 1; v < nv; ++v) w[c * nv + v] = inv * q[v];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kAuxBlock) rk_stage_kernel(const T* u, const T* us,
                                                             const T* div, T* out, T* us_out,
                                                             size_t n, double dt, StageCoef k) {
  const T sdt = T(k.beta * dt);
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    T s = us[c];
    if (k.has_delta) s = s + T(k.delta) * u[c];
    if (k.blended)
      out[c] = s + (T(k.gam0) * (u[c] - s) - sdt * div[c]);
    else
      out[c] = (T(k.gam0) * u[c] + T(k.gam1) * s) - sdt * div[c];
    if (us_out) us_out[c] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(kAuxBlock) flux_divergence_kernel(const T* fx, const T* fy,
                                                                    T* out, int nx, int ny,
                                                                    int nv, T dx, T dy) {
  const size_t n = (size_t)nx * ny * nv;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t c = k / nv;
    const int v = (int)(k % nv);
    const int j = (int)(c / nx), i = (int)(c % nx);
    const int ip = (i == nx - 1) ? 0 : i + 1, jp = (j == ny - 1) ? 0 : j + 1;
    const T ddx = dvd(fx[((size_t)j * nx + ip) * nv + v] - fx[k], dx);
    const T ddy = dvd(fy[((size_t)jp * nx + i) * nv + v] - fy[k], dy);
    out[k] = ddx + ddy;
  }
}

template <typename T>
__global__ void __launch_bounds__(kAuxBlock) stress_faces_kernel(const T* tau, T* fx, T* fy,
                                                                 int nx, int ny) {
  const size_t n = (size_t)nx * ny;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / nx), i = (int)(c % nx);
    const int im = (i == 0) ? nx - 1 : i - 1, jm = (j == 0) ? ny - 1 : j - 1;
    const T* t0 = tau + c * 3;
    const T* tw = tau + ((size_t)j * nx + im) * 3;
    const T* ts = tau + ((size_t)jm * nx + i) * 3;
    fx[c * 2 + 0] = T(0.5) * (tw[0] + t0[0]);
    fx[c * 2 + 1] = T(0.5) * (tw[1] + t0[1]);
    fy[c * 2 + 0] = T(0.5) * (ts[1] + t0[1]);
    fy[c * 2 + 1] = T(0.5) * (ts[2] + t0[2]);
  }
}

// Per-block float64 partial sums of each component (up to 8), fixed order.
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) box_sums_kernel(const T* u, size_t n, int nv,
                                                             double* partial) {
  __shared__ double s[8][kAuxBlock];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    for (int v = 0; v < nv; ++v) acc[v] += (double)u[c * nv + v];
  for (int v = 0; v < 8; ++v) s[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < 8; ++v) s[v][threadIdx.x] += s[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < nv; ++v) partial[(size_t)blockIdx.x * 8 + v] = s[v][0];
}

static __global__ void box_sums_finish_kernel(const double* partial, int nblocks, int nv,
                                              double* out) {
  if (threadIdx.x != 0) return;
  for (int v = 0; v < nv; ++v) {
    double a = 0.0;
    for (int b = 0; b < nblocks; ++b) a += partial[(size_t)b * 8 + v];
    out[v] = a;
  }
}

}  // namespace jx

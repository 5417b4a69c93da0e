// The fused RK-stage kernel: one launch per stage of jathenax/hydro.py:101-143.
//
// Decomposition.  A block of BX threads owns a strip of BX-6 output columns and `rows`
// output rows.  Thread t follows column x0+t-3 (periodic) and marches up the strip one
// row per iteration, so every input row is read from HBM once per strip:
//   * y direction: the thread keeps its column's last five primitive rows in registers
//     and builds the y-face states, the y Riemann flux and (with viscosity) the y-face
//     viscous flux without touching LDS for the stencil;
//   * x direction: the row being updated is staged in LDS; the thread reconstructs its
//     own cell, exchanges high edges (PPM4: interfaces) and face fluxes through LDS
//     (three barriers per row) and forms the x flux difference.  The 3+3 halo columns
//     are computed, not stored.
// Face states and fluxes never leave the CU.  The update applies the RK row
// (hydro.py:39-51), drag with the incoming primitives (hydro.py:122-125), the density
// floor (hydro.py:54-57) and, in the epilogue, either the CFL min for the next timestep
// (hydro.py:66-68) or the first moments of the forcing normalisation (forcing.py:141).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "jx_loop.hpp"
#include "jx_ops.hpp"

namespace jx {

constexpr int kMaxBlock = 256;   // threads per stage block (launch bound)

template <typename T> struct StageArgs {
  const T* __restrict__ u_in;   // stage input, read with halos (never aliases u_out)
  const T* u_sv;                // saved register, read pointwise (may alias u_out/s_out)
  T* u_out;                     // stage output
  T* s_out;                     // RK4 saved register output (nullptr when no delta)
  const T* __restrict__ raw;    // raw force (ny, nx, 2) when this stage feeds the kick
  ForceEval<T> force;           // FP32 viscous kernels: force.pq set, the force is evaluated
                                // per cell instead of read from raw
  const T* __restrict__ tau;    // learned stress (ny, nx, 3) of this stage, or nullptr
                                // (VISC kernels only)
  const Ctrl* ctrl;             // dt of the step and the done flag
  unsigned long long* dtkey;    // CFL-limit min slot (nullptr: no epilogue)
  double* partial;              // kick moments per block [3] (nullptr: no kick)
  int nx, ny, rows;
  int stage1;                   // saved register is the (floored) stage input itself
  int floor_saved;              // apply the density floor when reading u_sv
  int has_delta, blended, drag_on, smag_on;
  int visc_terms;               // VISC kernels: add the viscous flux (off: stress only)
  int exact;                    // dx, dy, cs are powers of two
  T dx, dy, cs, cs2, inv_cs, dfloor, eps, four_dx, four_dy, visc, smag_c, drag;
  T inv_dx, inv_dy, inv_four_dx, inv_four_dy;
  double gam0, gam1, beta, delta;
};

template <typename T>
__device__ __forceinline__ void block_reduce_store(double v0, double v1, double v2,
                                                   double* sred, double* out) {
  // Fixed butterfly inside each wave, then waves in index order: deterministic.
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off);
    v1 += __shfl_xor(v1, off);
    v2 += __shfl_xor(v2, off);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sred[3 * wave + 0] = v0;
    sred[3 * wave + 1] = v1;
    sred[3 * wave + 2] = v2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) {
      a += sred[3 * w + 0];
      b += sred[3 * w + 1];
      c += sred[3 * w + 2];
    }
    out[0] = a;
    out[1] = b;
    out[2] = c;
  }
}

// The same for N values (out[0 .. N-1]); sred holds N doubles per wave.
template <int N>
__device__ __forceinline__ void block_reduce_store_n(double (&v)[N], double* sred, double* out) {
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] += __shfl_xor(v[k], off);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) sred[N * wave + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double a = 0;
      for (int w = 0; w < nw; ++w) a += sred[N * w + k];
      out[k] = a;
    }
  }
}

// Kick sums per block written by a final stage that feeds the kick.  FP64: rho, rho r
// (the mean force; the second moments take another pass over the state, in the reference's
// form).  FP32 viscous kernels (the driven-turbulence configuration): also rho |r|^2, m.r
// and m, from which the second moments follow in one pass (forcing.py:141-145 expanded
// about the mean, summed in float64).  The inviscid FP32 kernels keep three sums: four more
// float64 accumulators would cost them a wave of occupancy.
template <typename T, bool VISC> constexpr int kick_sums() {
  return (Fast<T>::value && VISC) ? 7 : 3;
}

__device__ __forceinline__ void block_min_key(unsigned long long k, unsigned long long* sred,
                                              unsigned long long* slot) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off);
    k = o < k ? o : k;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sred[wave] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    unsigned long long m = sred[0];
    for (int w = 1; w < nw; ++w) m = sred[w] < m ? sred[w] : m;
    // The slot only decreases, so a stale read can only cause a redundant atomic.
    if (m < *reinterpret_cast<volatile unsigned long long*>(slot)) atomicMin(slot, m);
  }
}

// LDS layout of the stage kernel.  Each region holds one row of the block's BX columns
// plus two pad columns on either side, with a column's values interleaved at an odd word
// stride (lds_stride): neighbouring lanes then fall in different banks, and every access
// is one base register per region plus a compile-time offset (column shift times the
// stride plus the variable).  Regions: primitives, high edges, face fluxes, PPM4
// interfaces (lds_stride words per column each), then three rows of (rho, vx, vy) for the
// viscous faces, then 64 doubles for the block reductions.
__host__ __device__ constexpr int lds_stride(int nv) { return nv | 1; }
template <typename T>
__host__ __device__ constexpr size_t stage_smem_bytes(int nv, bool visc, int bx) {
  return (size_t)(4 * lds_stride(nv) + (visc ? 9 : 0)) * (bx + 4) * sizeof(T) +
         64 * sizeof(double);
}
// Pad columns on either side of each region.  They are set to 1 once per block: only
// halo lanes, whose results are never used, read them, and 1 keeps their values finite.
constexpr int kLdsPad = 2;

template <int P> using IC = std::integral_constant<int, P>;

// ST1: the first stage of a step (its RK row is (0, 1) and the saved register is the stage
// input itself), which reads no saved register.
template <typename T, int RECON, int RS, int NV, bool VISC, bool EXACT, bool ST1>
__global__ void __launch_bounds__(kMaxBlock)
stage_kernel(StageArgs<T> args) {
  // A local copy keeps the lambda below from taking the kernel argument's address.
  const StageArgs<T> a = args;
  if (a.ctrl->done) return;
  // PPM4 shares each y interface between the two cells it bounds, so its register ring
  // holds rows r-3..r; the other reconstructions need the full 5-row stencil.
  constexpr bool SHARED = RECON == RECON_PPM4;
  constexpr int R = SHARED ? 4 : 5;
  extern __shared__ double smem_d[];
  const int BX = blockDim.x;
  const int t = threadIdx.x;
  // This lane's column in each LDS region (see stage_smem_bytes); column shift d and
  // variable v of a region are at [d * NS + v].
  constexpr int NS = lds_stride(NV);
  const int LW = (BX + 2 * kLdsPad) * NS;   // words per region
  T* const lds = reinterpret_cast<T*>(smem_d);
  T* const sW = lds + (t + kLdsPad) * NS;   // primitives of the row being updated
  T* const sH = sW + LW;                    // x high edges
  T* const sF = sH + LW;                    // x face fluxes
  T* const sI = sF + LW;                    // PPM4 x interfaces
  // viscous rows: slot k (row q with q mod 3 == k) at sV + k * VW, (rho, vx, vy) per column
  const int VW = (BX + 2 * kLdsPad) * 3;
  T* const sV = lds + 4 * LW + (t + kLdsPad) * 3;
  double* sred = reinterpret_cast<double*>(lds + 4 * LW + (VISC ? 3 * VW : 0));
  {
    // Pad columns: two on either side of every region, written once (lanes 0-3).
    if (t < 2 * kLdsPad) {
      const int c = (t < kLdsPad) ? t : BX + t;
      for (int k = 0; k < 4; ++k)
        for (int v = 0; v < NS; ++v) lds[k * LW + c * NS + v] = T(1);
      if constexpr (VISC)
        for (int k = 0; k < 3; ++k)
          for (int v = 0; v < 3; ++v) lds[4 * LW + k * VW + c * 3 + v] = T(1);
    }
  }

  const int nx = a.nx, ny = a.ny;
  const int OX = BX - 6;
  // XCD-aware tile order.  Workgroups go to the 8 XCDs round-robin by linear id, and each
  // XCD has its own L2.  Within each window of 8 bands (8 x gridDim.x consecutive ids), the
  // ids XCD x receives take every strip of band 8k + x, so horizontally adjacent strips of
  // a band run on one XCD at the same time and the cache lines they share (their 6-column
  // overlap) are fetched into that L2 once.  Windows keep the band order, so blocks beyond
  // the resident ones still take the last bands; a final partial window keeps the plain
  // order.
  int strip, band;
  {
    const int W = gridDim.x;
    const int lin = blockIdx.x + W * blockIdx.y;
    const int win = lin / (8 * W), within = lin - win * 8 * W;
    if (8 * (win + 1) <= (int)gridDim.y) {
      band = 8 * win + (within & 7);
      strip = within >> 3;
    } else {
      band = 8 * win + within / W;
      strip = within - (within / W) * W;
    }
  }
  const int x0 = strip * OX;
  int xc = (x0 + t - 3) % nx;
  if (xc < 0) xc += nx;
  const bool col_out = (t >= 3) && (t < BX - 3) && (x0 + t - 3 < nx);
  // column shifts in LDS words
  constexpr int M2 = -2 * NS, M1 = -NS, P1 = NS, P2 = 2 * NS;
  const int y0 = band * a.rows;
  const int yend = min(y0 + a.rows, ny);
  const int nit = (yend - y0) + 6;

  const double dt = a.ctrl->dt;
  const T sdt = T(a.beta * dt);
  const Gas<T> gas{a.cs, a.cs2, a.inv_cs};
  const Visc<T> vp{a.visc, a.smag_c, a.smag_on != 0};
  const Over<T, EXACT> over_dx{a.dx, a.inv_dx}, over_dy{a.dy, a.inv_dy};
  const Over<T, EXACT> over_4dx{a.four_dx, a.inv_four_dx}, over_4dy{a.four_dy, a.inv_four_dy};
  const T dfloor = a.dfloor;
  const T drag_coef = sdt * a.drag;
  // FP32 Roe/LLF fluxes (and the viscous and stress terms added to them) are doubled; the
  // update's flux coefficients take the halving.
  constexpr bool X2 = flux_x2<RS, T>();
  const T cx = flux_coef<T>(a.beta * dt, a.inv_dx, X2);
  const T cy = flux_coef<T>(a.beta * dt, a.inv_dy, X2);

  T w[R][NV];   // primitives, row q in slot (q - y0 + 3) mod R
  T u[R][NV];   // floored conserved state, same slots
  T ip[NV];     // PPM4: clamped interface below row r-2
  T hy[NV];     // y high edge of row r-3
  T gp[NV];     // y flux at the low face of row r-3
  T gn[NV];     // y flux at the low face of row r-2
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    ip[v] = hy[v] = gp[v] = gn[v] = T(0);
#pragma unroll
    for (int k = 0; k < R; ++k) w[k][v] = u[k][v] = T(0);
  }
  constexpr int KS = kick_sums<T, VISC>();
  double acc[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) acc[k] = 0.0;
  // Kick moments from the evaluated force: this lane's column factors of each mode group.
  ForceCols<T> fcol{};
  if constexpr (KS == 7) {
    if (a.partial && a.force.pq) fcol = force_cols(a.force, nx, xc);
  }
  CflMin<T> cfl(a.dx, a.dy, a.cs);
  // Loads run ahead of their use so that HBM latency under full load (longer than one row
  // iteration) stays hidden.  With the 4-row ring (even, so the loop parity is known at
  // compile time) the input row is fetched two iterations ahead and the saved register
  // one iteration ahead, each into a two-slot register ring indexed by parity (no register
  // moves); with the 5-row ring, one row ahead and in the same iteration.
  constexpr bool RING2 = (R % 2) == 0;
  T pfq[2][NV];   // prefetched input rows: slot (i & 1) holds row r (RING2), pfq[0] row r+1
  T svq[2][NV];   // saved register rows: iteration i updates from slot ((i + 1) & 1)
  const size_t row_len = (size_t)nx * NV;
  const int row_bytes = (int)(row_len * sizeof(T));
  // Rows are read and written through per-row buffer resources (row_load/row_store) at
  // these 32-bit lane byte offsets.  lane_out wraps for the left halo lanes, which never
  // store.
  const unsigned lane_in = (unsigned)(xc * NV) * (unsigned)sizeof(T);
  const unsigned lane_out = (unsigned)((x0 + t - 3) * NV) * (unsigned)sizeof(T);
  const T* const sv_base = a.u_sv;
  {
    int yn = (y0 - 3) % ny;
    if (yn < 0) yn += ny;
    row_load<T, NV>(row_rsrc(a.u_in + (size_t)yn * row_len, row_bytes), lane_in, pfq[0]);
    if constexpr (RING2) {
      yn = (y0 - 2) % ny;
      if (yn < 0) yn += ny;
      row_load<T, NV>(row_rsrc(a.u_in + (size_t)yn * row_len, row_bytes), lane_in, pfq[1]);
      // iteration 0 updates nothing: any row will do (the block's first, which stays cached)
      if constexpr (!ST1)
        row_load<T, NV>(row_rsrc(sv_base + (size_t)y0 * row_len, row_bytes), lane_in, svq[1]);
    }
  }

  // WARM: one of the first six iterations, where the row ring, the y faces and the x work
  // start up in turn; in the steady iterations (i >= 6) every past(k) test folds to true.
  auto body = [&](auto pc, auto warm, int i) {
    constexpr int P = decltype(pc)::value;
    constexpr bool WARM = decltype(warm)::value;
    auto past = [&](int k) { return !WARM || i >= k; };
    constexpr int S0 = P, S1 = (P + R - 1) % R, S2 = (P + R - 2) % R, S3 = (P + R - 3) % R;
    constexpr int S4 = (P + R - 4) % R;   // row r-4 (5-row ring only)
    const int r = y0 - 3 + i;
    // Both loads of an iteration are unconditional, so the memory-wait pass sees
    // straight-line code and waits for each load alone.  Where an iteration needs no new
    // row, it re-reads one that is still cached: the saved register during the warm-up
    // (the block's first row, read by the input stream at the same time) and in the final
    // iteration (row r-3, loaded the iteration before; row r-2 = yend may be row ny);
    // the input stream past the block's top halo (row r, loaded two iterations ago).
    constexpr int PA = P & 1, PB = (P + 1) & 1;
    T svr[NV];
    if constexpr (!ST1) {
      if constexpr (RING2) {
        // row r-3 arrived last iteration; row r-2 is the next iteration's
#pragma unroll
        for (int v = 0; v < NV; ++v) svr[v] = svq[PB][v];
        const int q = (past(5) && r - 2 < yend) ? r - 2 : (past(6) ? r - 3 : y0);
        const T* sv = sv_base + (size_t)q * row_len;
        row_load<T, NV>(row_rsrc(sv, row_bytes), lane_in, svq[PA]);
      } else {
        const T* sv = sv_base + (size_t)(past(6) ? r - 3 : y0) * row_len;
        row_load<T, NV>(row_rsrc(sv, row_bytes), lane_in, svr);
      }
    }
    {
      // RING2: row r was fetched two iterations ago into this slot; issue row r+2 into
      // it.  Otherwise row r came one iteration ago; issue row r+1.  Past the end, the
      // wrapped row (only the 3-row halos wrap).
      constexpr int PQ = RING2 ? PA : 0;
#pragma unroll
      for (int v = 0; v < NV; ++v) u[S0][v] = pfq[PQ][v];
      // rows up to yend + 2 are needed
      int yn = (r + (RING2 ? 2 : 1) <= yend + 2) ? r + (RING2 ? 2 : 1) : r;
      if (yn < 0 || yn >= ny) {
        yn %= ny;
        if (yn < 0) yn += ny;
      }
      row_load<T, NV>(row_rsrc(a.u_in + (size_t)yn * row_len, row_bytes), lane_in, pfq[PQ]);
    }
    {
      // recover_isothermal (hydro.py:54-60): floor, then (1/rho) * momentum.
      const T rho = floor_to(u[S0][0], dfloor);
      u[S0][0] = rho;
      const T inv = dvd(T(1), rho);
      w[S0][0] = rho;
#pragma unroll
      for (int v = 1; v < NV; ++v) w[S0][v] = inv * u[S0][v];
    }
    if constexpr (VISC) {
      if (past(2)) {
        T* slot = sV + ((r - 2 + 3 * ny + 3) % 3) * VW;
        slot[0] = w[S2][0];
        slot[1] = w[S2][1];
        slot[2] = w[S2][2];
      }
    }
    if (past(6)) {
#pragma unroll
      for (int v = 0; v < NV; ++v) sW[v] = w[S3][v];
    }
    if (VISC ? (past(2)) : (past(6))) __syncthreads();   // B1
    // The x stencil of row r-3 is read now, so that the LDS latency overlaps the
    // y-direction arithmetic below.
    T xm2[NV], xm1[NV], xp1[NV], xp2[NV];
    {   // unconditional: before the first barrier these are in-bounds, unused LDS words
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        xm2[v] = sW[M2 + v];
        xm1[v] = sW[M1 + v];
        xp1[v] = sW[P1 + v];
        if constexpr (!SHARED) xp2[v] = sW[P2 + v];
      }
    }

    // ---- y direction: edges of row r-2, flux at its low face -------------------------
    auto y_face = [&](T (&lo)[NV], T (&hi)[NV]) {
      if (past(5)) {
        // y faces: normal = vy (index 2), tangent = vx (index 1) (fluxes.py:209-214)
        T f0, f1, f2;
        riemann_x<RS, EXACT, X2>(hy[0], lo[0], hy[2], lo[2], hy[1], lo[1], gas, f0, f1, f2);
        gn[0] = f0;
        gn[1] = f2;
        gn[2] = f1;
#pragma unroll
        for (int v = 3; v < NV; ++v) gn[v] = f0 * ((f0 >= T(0)) ? hy[v] : lo[v]);
        if (VISC && a.visc_terms) {
          // low y face of row s = r-2, column xc (fluxes.py:227-238 with axis=0)
          const T* rs = sV + ((r - 2 + 3 * ny + 3) % 3) * VW;   // row s
          const T* rm = sV + ((r - 3 + 3 * ny + 3) % 3) * VW;   // row s-1
          const T uy = over_dy(w[S2][1] - w[S3][1]);     // normal derivatives
          const T vy = over_dy(w[S2][2] - w[S3][2]);
          const T ux = over_4dx((rs[3 + 1] + rm[3 + 1]) - (rs[-3 + 1] + rm[-3 + 1]));
          const T vx = over_4dx((rs[3 + 2] + rm[3 + 2]) - (rs[-3 + 2] + rm[-3 + 2]));
          T fmx_, fmy_;
          visc_face<0, X2>(ux, vy, uy, vx, w[S2][0] + w[S3][0], vp, fmx_, fmy_);
          gn[1] = gn[1] + fmx_;
          gn[2] = gn[2] + fmy_;
        }
        if (VISC && a.tau) {
          // Learned stress averaged onto the low y face of row r-2, added after the
          // viscous flux (hydro.py:111-119, fluxes.py:269-280): (tau_xy, tau_yy).
          int s0 = (r - 2) % ny, sm = (r - 3) % ny;
          if (s0 < 0) s0 += ny;
          if (sm < 0) sm += ny;
          const T* t0 = a.tau + (size_t)s0 * nx * 3 + xc * 3;
          const T* tm = a.tau + (size_t)sm * nx * 3 + xc * 3;
          if constexpr (X2) {
            gn[1] = gn[1] + (tm[1] + t0[1]);
            gn[2] = gn[2] + (tm[2] + t0[2]);
          } else {
            gn[1] = gn[1] + T(0.5) * (tm[1] + t0[1]);
            gn[2] = gn[2] + T(0.5) * (tm[2] + t0[2]);
          }
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) hy[v] = hi[v];
    };
    if constexpr (SHARED) {
      if (past(3)) {
        T in[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
          in[v] = ppm4_interface(w[S3][v], w[S2][v], w[S1][v], w[S0][v]);   // below row r-1
        if (past(4)) {
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
      if (past(4)) {
        T lo[NV], hi[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
          cell_edges<RECON>(w[S4][v], w[S3][v], w[S2][v], w[S1][v], w[S0][v], a.eps, lo[v],
                            hi[v]);
        if constexpr (floors_faces<RECON>()) {
          lo[0] = fmx(lo[0], dfloor);
          hi[0] = fmx(hi[0], dfloor);
        }
        y_face(lo, hi);
      }
    }

    // ---- x direction for row j = r-3, then the update --------------------------------
    if (past(6)) {
      T xl[NV] = {}, xh[NV] = {};
      T l[NV] = {};   // left state at the cell's low x face: the left neighbour's high edge
      if constexpr (SHARED) {
        // Each lane evaluates the clamped interface at its cell's low face once; the
        // high edge is the next lane's interface (same value the reference computes
        // twice, see ppm4_interface). The lane also limits its left neighbour's edges
        // (the same operation on the same values as that lane), which gives the left
        // state without another exchange.
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          xl[v] = ppm4_interface(xm2[v], xm1[v], w[S3][v], xp1[v]);
          sI[v] = xl[v];
        }
        __syncthreads();   // B1b
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          T nlo = sI[M1 + v], nhi = xl[v];
          ppm4_limit(xm1[v], nlo, nhi);
          l[v] = nhi;
          xh[v] = sI[P1 + v];
          ppm4_limit(w[S3][v], xl[v], xh[v]);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          cell_edges<RECON>(xm2[v], xm1[v], w[S3][v], xp1[v], xp2[v], a.eps, xl[v], xh[v]);
        if constexpr (floors_faces<RECON>()) {
          xl[0] = fmx(xl[0], dfloor);
          xh[0] = fmx(xh[0], dfloor);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sH[v] = xh[v];
        __syncthreads();   // B2
#pragma unroll
        for (int v = 0; v < NV; ++v) l[v] = sH[M1 + v];
      }
      T fx[NV] = {};
      {
        T f0, f1, f2;
        riemann_x<RS, EXACT, X2>(l[0], xl[0], l[1], xl[1], l[2], xl[2], gas, f0, f1, f2);
        fx[0] = f0;
        fx[1] = f1;
        fx[2] = f2;
#pragma unroll
        for (int v = 3; v < NV; ++v) fx[v] = f0 * ((f0 >= T(0)) ? l[v] : xl[v]);
        if (VISC && a.visc_terms) {
          // low x face of row j = r-3 at column xc (fluxes.py:227-238 with axis=1)
          const T* rj = sV + ((r - 3 + 3 * ny + 3) % 3) * VW;   // row j
          const T* rp = sV + ((r - 2 + 3 * ny + 3) % 3) * VW;   // row j+1
          const T* rq = sV + ((r - 4 + 3 * ny + 3) % 3) * VW;   // row j-1
          const T ux = over_dx(rj[1] - rj[-3 + 1]);
          const T vx = over_dx(rj[2] - rj[-3 + 2]);
          const T uy = over_4dy((rp[1] + rp[-3 + 1]) - (rq[1] + rq[-3 + 1]));
          const T vy = over_4dy((rp[2] + rp[-3 + 2]) - (rq[2] + rq[-3 + 2]));
          T fmx_, fmy_;
          visc_face<1, X2>(ux, vy, uy, vx, rj[0] + rj[-3], vp, fmx_, fmy_);
          fx[1] = fx[1] + fmx_;
          fx[2] = fx[2] + fmy_;
        }
        if (VISC && a.tau) {
          // Learned stress on the low x face of column xc of row r-3: (tau_xx, tau_xy).
          const int xm = (xc == 0) ? nx - 1 : xc - 1;
          const T* trow = a.tau + (size_t)(r - 3) * nx * 3;
          if constexpr (X2) {
            fx[1] = fx[1] + (trow[xm * 3 + 0] + trow[xc * 3 + 0]);
            fx[2] = fx[2] + (trow[xm * 3 + 1] + trow[xc * 3 + 1]);
          } else {
            fx[1] = fx[1] + T(0.5) * (trow[xm * 3 + 0] + trow[xc * 3 + 0]);
            fx[2] = fx[2] + T(0.5) * (trow[xm * 3 + 1] + trow[xc * 3 + 1]);
          }
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) sF[v] = fx[v];
      __syncthreads();   // B3
      if (col_out) {
        const int j = r - 3;
        T* const out_row = a.u_out + (size_t)j * row_len;
        T un[NV] = {};
        T us[NV] = {};
        if constexpr (ST1) {
#pragma unroll
          for (int v = 0; v < NV; ++v) us[v] = u[S3][v];
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) us[v] = svr[v];
          if (a.floor_saved) us[0] = floor_to(us[0], dfloor);
        }
        if (a.has_delta) {
          // RK4 saved-register accumulation (hydro.py:42-43)
#pragma unroll
          for (int v = 0; v < NV; ++v) us[v] = us[v] + T(a.delta) * u[S3][v];
          row_store<T, NV, kNonTemporal>(row_rsrc(a.s_out + (size_t)j * row_len, row_bytes), lane_out, us);
        }
        if constexpr (Fast<T>::value) {
          T dF[NV], dG[NV];
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            dF[v] = sF[P1 + v] - fx[v];
            dG[v] = gn[v] - gp[v];
          }
          rk_update_fast<NV>(u[S3], us, dF, dG, ST1, a.blended != 0, T(a.gam0),
                             T(a.gam1), cx, cy, un);
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            const T div = over_dx(sF[P1 + v] - fx[v]) + over_dy(gn[v] - gp[v]);
            if (a.blended)
              un[v] = us[v] + (T(a.gam0) * (u[S3][v] - us[v]) - sdt * div);
            else
              un[v] = (T(a.gam0) * u[S3][v] + T(a.gam1) * us[v]) - sdt * div;
          }
        }
        if (a.drag_on) {
          // drag with the stage's incoming primitives (hydro.py:122-125); the empty
          // volatile asm keeps this a branch rather than selects every zone would execute
          __asm__ volatile("");
          const T c = drag_coef * w[S3][0];
          un[1] = un[1] + (-(c * w[S3][1]));
          un[2] = un[2] + (-(c * w[S3][2]));
        }
        if (a.partial) {
          // The kick normalises against the post-drag state; the floor comes after it.
          if constexpr (KS == 7) {
            T f0, f1;
            if (a.force.pq) {
              force_at(a.force, fcol, j, f0, f1);
            } else {
              const T* rf = a.raw + (size_t)j * nx * 2 + (x0 + t - 3) * 2;
              f0 = rf[0];
              f1 = rf[1];
            }
            // float products are exact in float64
            const double rho = un[0], r0 = f0, r1 = f1;
            const double m0 = un[1], m1 = un[2];
            acc[0] += rho;
            acc[1] += rho * r0;
            acc[2] += rho * r1;
            acc[3] += rho * (r0 * r0 + r1 * r1);
            acc[4] += m0 * r0 + m1 * r1;
            acc[5] += m0;
            acc[6] += m1;
          } else {
            const T* rf = a.raw + (size_t)j * nx * 2 + (x0 + t - 3) * 2;
            acc[0] += (double)un[0];
            acc[1] += (double)(un[0] * rf[0]);
            acc[2] += (double)(un[0] * rf[1]);
          }
        } else {
          un[0] = floor_to(un[0], dfloor);
          if (a.dtkey) {
            const T inv = dvd(T(1), un[0]);
            cfl.add(inv * un[1], inv * un[2]);
          }
        }
        row_store<T, NV, kNonTemporal>(row_rsrc(out_row, row_bytes), lane_out, un);
      }
    }
    // gn is set from iteration 5 on and gp is first read at iteration 6, so the copy can
    // be unconditional.
#pragma unroll
    for (int v = 0; v < NV; ++v) gp[v] = gn[v];
  };

  // nit = rows + 6 >= 7: the six warm-up iterations always run.
  using Warm = std::true_type;
  using Steady = std::false_type;
  body(IC<0>{}, Warm{}, 0);
  body(IC<1>{}, Warm{}, 1);
  body(IC<2>{}, Warm{}, 2);
  body(IC<3>{}, Warm{}, 3);
  body(IC<4 % R>{}, Warm{}, 4);
  body(IC<5 % R>{}, Warm{}, 5);
  auto round = [&](auto k0, int i) {
    constexpr int K = decltype(k0)::value;
    body(IC<(6 + K) % R>{}, Steady{}, i + K);
  };
  int i0 = 6;
  constexpr bool WHOLE = Fast<T>::value && NV == 3 && !VISC && RECON != RECON_PPMX &&
                        RECON != RECON_WENOZ;
  if constexpr (WHOLE) {
    // FP32 hydro without viscosity (DC, PLM, PPM4): the steady loop runs whole rounds of two
    // rings' worth of rows (8 or 10), with no exit between them.  An exit after a row makes the ring and prefetch registers live-out
    // there, and the copies that merge them wait for every load in flight; the round's back
    // edge is the only such point, once per 2R rows.  The last (nit - 6) mod 2R rows follow.
    constexpr int RR = 2 * R;
    for (; i0 + RR <= nit; i0 += RR) {
      round(IC<0>{}, i0); round(IC<1>{}, i0); round(IC<2>{}, i0); round(IC<3>{}, i0);
      round(IC<4>{}, i0); round(IC<5>{}, i0); round(IC<6>{}, i0); round(IC<7>{}, i0);
      if constexpr (R == 5) { round(IC<8>{}, i0); round(IC<9>{}, i0); }
    }
    if (i0 + 0 < nit) round(IC<0>{}, i0);
    if (i0 + 1 < nit) round(IC<1>{}, i0);
    if (i0 + 2 < nit) round(IC<2>{}, i0);
    if (i0 + 3 < nit) round(IC<3>{}, i0);
    if (i0 + 4 < nit) round(IC<4>{}, i0);
    if (i0 + 5 < nit) round(IC<5>{}, i0);
    if (i0 + 6 < nit) round(IC<6>{}, i0);
    if constexpr (R == 5) {
      if (i0 + 7 < nit) round(IC<7>{}, i0);
      if (i0 + 8 < nit) round(IC<8>{}, i0);
    }
  } else {
    // Rounds of one ring with a test before each row.  For FP64 (twice the registers per
    // value), scalars, the viscous kernels and the PPMX / WENO-Z stencils the whole-round
    // form makes the compiler keep part of the state in scratch memory.
    for (; i0 < nit; i0 += R) {
      round(IC<0>{}, i0);
      if (i0 + 1 < nit) round(IC<1>{}, i0);
      if (i0 + 2 < nit) round(IC<2>{}, i0);
      if (i0 + 3 < nit) round(IC<3>{}, i0);
      if constexpr (R == 5) {
        if (i0 + 4 < nit) round(IC<4>{}, i0);
      }
    }
  }

  if (a.partial) {
    block_reduce_store_n<KS>(acc, sred,
                             a.partial + KS * (size_t)(band * gridDim.x + strip));
  } else if (a.dtkey) {
    block_min_key(cfl.key(), reinterpret_cast<unsigned long long*>(sred), a.dtkey);
  }
}

// One launch.  With events (timing), the start and stop timestamps are taken by the
// dispatch itself (hipExtLaunchKernel), so timing adds no packets between kernels.
template <typename T, typename K>
hipError_t launch_one(K kernel, dim3 grid, int bx, size_t smem, hipStream_t st,
                      const StageArgs<T>& a, hipEvent_t ev0, hipEvent_t ev1) {
  if (ev0) {
    void* args[] = {const_cast<StageArgs<T>*>(&a)};
    return hipExtLaunchKernel(reinterpret_cast<const void*>(kernel), grid, dim3(bx), args, smem,
                              st, ev0, ev1, 0);
  }
  hipLaunchKernelGGL(kernel, grid, dim3(bx), smem, st, a);
  return hipGetLastError();
}

// Host-side launch for one (precision, reconstruction) pair; the Riemann solver, the
// number of variables and the viscous switch are dispatched here.
template <typename T, int RECON, bool EXACT>
hipError_t launch_stage_e(int rs, int nv, bool visc, const StageArgs<T>& a, dim3 grid, int bx,
                          hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
#define JX_LAUNCH(RS_, NV_, V_)                                                             \
  do {                                                                                     \
    if (a.stage1)                                                                          \
      return launch_one<T>(stage_kernel<T, RECON, RS_, NV_, V_, EXACT, true>, grid, bx,     \
                           stage_smem_bytes<T>(NV_, V_, bx), st, a, ev0, ev1);             \
    return launch_one<T>(stage_kernel<T, RECON, RS_, NV_, V_, EXACT, false>, grid, bx,      \
                         stage_smem_bytes<T>(NV_, V_, bx), st, a, ev0, ev1);               \
  } while (0)
#define JX_NV(RS_)                                                                          \
  do {                                                                                     \
    if (nv == 3) { if (visc) JX_LAUNCH(RS_, 3, true); else JX_LAUNCH(RS_, 3, false); }     \
    if (nv == 4) { if (visc) JX_LAUNCH(RS_, 4, true); else JX_LAUNCH(RS_, 4, false); }     \
    if (nv == 5) { if (visc) JX_LAUNCH(RS_, 5, true); else JX_LAUNCH(RS_, 5, false); }     \
  } while (0)
  if (rs == RS_LLF) JX_NV(RS_LLF);
  if (rs == RS_HLLE) JX_NV(RS_HLLE);
  if (rs == RS_ROE) JX_NV(RS_ROE);
#undef JX_NV
#undef JX_LAUNCH
  return hipErrorInvalidValue;
}

// Resident blocks per CU for the kernel a launch would use (0 on error).
template <typename T, int RECON, bool EXACT>
int occupancy_e(int rs, int nv, bool visc, int bx) {
  int nb = 0;
#define JX_OCC(RS_, NV_, V_)                                                                \
  do {                                                                                     \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                       \
            &nb, stage_kernel<T, RECON, RS_, NV_, V_, EXACT, false>, bx,                   \
            stage_smem_bytes<T>(NV_, V_, bx)) != hipSuccess)                               \
      nb = 0;                                                                              \
    return nb;                                                                             \
  } while (0)
#define JX_ONV(RS_)                                                                         \
  do {                                                                                     \
    if (nv == 3) { if (visc) JX_OCC(RS_, 3, true); else JX_OCC(RS_, 3, false); }           \
    if (nv == 4) { if (visc) JX_OCC(RS_, 4, true); else JX_OCC(RS_, 4, false); }           \
    if (nv == 5) { if (visc) JX_OCC(RS_, 5, true); else JX_OCC(RS_, 5, false); }           \
  } while (0)
  if (rs == RS_LLF) JX_ONV(RS_LLF);
  if (rs == RS_HLLE) JX_ONV(RS_HLLE);
  if (rs == RS_ROE) JX_ONV(RS_ROE);
#undef JX_ONV
#undef JX_OCC
  return 0;
}

template <typename T, int RECON>
int stage_occupancy(int rs, int nv, bool visc, int bx, bool exact) {
  if constexpr (Fast<T>::value) {
    return occupancy_e<T, RECON, true>(rs, nv, visc, bx);
  } else {
    return exact ? occupancy_e<T, RECON, true>(rs, nv, visc, bx)
                 : occupancy_e<T, RECON, false>(rs, nv, visc, bx);
  }
}

// FP32 always multiplies by reciprocals; FP64 does so only when dx, dy and cs are powers
// of two (a.exact), where the product equals the IEEE quotient.
template <typename T, int RECON>
hipError_t launch_stage(int rs, int nv, bool visc, const StageArgs<T>& a, dim3 grid, int bx,
                        hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
  if constexpr (Fast<T>::value) {
    return launch_stage_e<T, RECON, true>(rs, nv, visc, a, grid, bx, st, ev0, ev1);
  } else {
    if (a.exact) return launch_stage_e<T, RECON, true>(rs, nv, visc, a, grid, bx, st, ev0, ev1);
    return launch_stage_e<T, RECON, false>(rs, nv, visc, a, grid, bx, st, ev0, ev1);
  }
}

}  // namespace jx

// Each stage_<prec>_<recon>.hip translation unit instantiates one launcher and its
// occupancy query.
#define JX_STAGE_LAUNCHER(T, RECON, NAME)                                                   \
  hipError_t NAME(int rs, int nv, bool visc, const jx::StageArgs<T>& a, dim3 grid, int bx,  \
                  hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {                        \
    return jx::launch_stage<T, RECON>(rs, nv, visc, a, grid, bx, st, ev0, ev1);            \
  }                                                                                        \
  int NAME##_occupancy(int rs, int nv, bool visc, int bx, bool exact) {                    \
    return jx::stage_occupancy<T, RECON>(rs, nv, visc, bx, exact);                         \
  }

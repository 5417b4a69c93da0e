// Unfused stage operators: the traced (debug) stage path and the learned-stress path.
//
// Each kernel restates one array function of the reference over the whole grid and writes
// its result to HBM, using the same scalar operators as the fused stage kernel
// (jx_ops.hpp), so a traced stage reproduces the fused stage bit for bit while exposing
// the face states and fluxes the fused kernel keeps on chip.  Grid-stride loops, one
// thread per cell (or face).
#pragma once
#include <hip/hip_runtime.h>

#include "jx_aux.hpp"
#include "jx_ops.hpp"

namespace jx {

// Periodic neighbour `o` cells away along `axis` (1 = x, 0 = y) of cell (j, i).
__device__ __forceinline__ size_t along(int j, int i, int o, int axis, int nx, int ny) {
  if (axis == 1) {
    int ii = (i + o) % nx;
    if (ii < 0) ii += nx;
    return (size_t)j * nx + ii;
  }
  int jj = (j + o) % ny;
  if (jj < 0) jj += ny;
  return (size_t)jj * nx + i;
}

// reconstruct_faces (fluxes.py:124-131): left = high edge of the cell below the face,
// right = low edge of the cell above it; PPMX and WENO-Z floor the face density unless
// dfloor is NaN (the registry functions *_face_states, fluxes.py:6-117, do not floor).
template <typename T, int RECON>
__global__ void __launch_bounds__(kAuxBlock) face_states_kernel(const T* W, T* left, T* right,
                                                                int nx, int ny, int nv,
                                                                int axis, T dfloor, T eps) {
  const size_t n = (size_t)nx * ny;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / nx), i = (int)(c % nx);
    size_t nb[6];
#pragma unroll
    for (int o = 0; o < 6; ++o) nb[o] = along(j, i, o - 3, axis, nx, ny) * nv;
    for (int v = 0; v < nv; ++v) {
      T q[6];
#pragma unroll
      for (int o = 0; o < 6; ++o) q[o] = W[nb[o] + v];
      T lo_b, hi_b, lo_c, hi_c;
      cell_edges<RECON>(q[0], q[1], q[2], q[3], q[4], eps, lo_b, hi_b);   // cell below
      cell_edges<RECON>(q[1], q[2], q[3], q[4], q[5], eps, lo_c, hi_c);   // this cell
      T l = hi_b, r = lo_c;
      if (floors_faces<RECON>() && v == 0 && dfloor == dfloor) {
        l = fmx(l, dfloor);
        r = fmx(r, dfloor);
      }
      left[c * nv + v] = l;
      right[c * nv + v] = r;
    }
  }
}

// isothermal_flux (fluxes.py:207-219): y faces take vy as the normal and swap the two
// momentum fluxes back to (x, y) order; scalars are the mass flux times the upwind value.
template <typename T, int RS>
__global__ void __launch_bounds__(kAuxBlock) riemann_faces_kernel(const T* left,
                                                                  const T* right, T* flux,
                                                                  size_t n, int nv, int axis,
                                                                  Gas<T> gas) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T* l = left + c * nv;
    const T* r = right + c * nv;
    T* f = flux + c * nv;
    T f0, f1, f2;
    if (axis == 1) {
      riemann<RS, false>(l[0], r[0], l[1], r[1], l[2], r[2], gas, f0, f1, f2);
      f[1] = f1;
      f[2] = f2;
    } else {
      riemann<RS, false>(l[0], r[0], l[2], r[2], l[1], r[1], gas, f0, f1, f2);
      f[1] = f2;
      f[2] = f1;
    }
    f[0] = f0;
    for (int v = 3; v < nv; ++v) f[v] = f0 * ((f0 >= T(0)) ? l[v] : r[v]);
  }
}

// viscous_momentum_flux (fluxes.py:227-266) at the low face of each cell along `axis`:
// out (ny, nx, 2) = -mu * stress.
// Divisions by dx, dy, 4dx, 4dy go through Over<T, false>: IEEE division in FP64 (equal to
// the fused kernel's, which multiplies only by exact power-of-two reciprocals), reciprocal
// multiplies in FP32 (as the fused kernel).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) viscous_faces_kernel(const T* W, T* out, int nx,
                                                                  int ny, int nv, int axis,
                                                                  Over<T, false> over_dx,
                                                                  Over<T, false> over_dy,
                                                                  Over<T, false> over_4dx,
                                                                  Over<T, false> over_4dy,
                                                                  Visc<T> vp) {
  const size_t n = (size_t)nx * ny;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / nx), i = (int)(c % nx);
    const int im = (i == 0) ? nx - 1 : i - 1, ip = (i == nx - 1) ? 0 : i + 1;
    const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
    auto w = [&](int jj, int ii, int v) { return W[((size_t)jj * nx + ii) * nv + v]; };
    T fmx_, fmy_;
    if (axis == 1) {
      const T ux = over_dx(w(j, i, 1) - w(j, im, 1));
      const T vx = over_dx(w(j, i, 2) - w(j, im, 2));
      const T uy = over_4dy((w(jp, i, 1) + w(jp, im, 1)) - (w(jm, i, 1) + w(jm, im, 1)));
      const T vy = over_4dy((w(jp, i, 2) + w(jp, im, 2)) - (w(jm, i, 2) + w(jm, im, 2)));
      visc_face<1>(ux, vy, uy, vx, w(j, i, 0) + w(j, im, 0), vp, fmx_, fmy_);
    } else {
      const T uy = over_dy(w(j, i, 1) - w(jm, i, 1));
      const T vy = over_dy(w(j, i, 2) - w(jm, i, 2));
      const T ux = over_4dx((w(j, ip, 1) + w(jm, ip, 1)) - (w(j, im, 1) + w(jm, im, 1)));
      const T vx = over_4dx((w(j, ip, 2) + w(jm, ip, 2)) - (w(j, im, 2) + w(jm, im, 2)));
      visc_face<0>(ux, vy, uy, vx, w(j, i, 0) + w(jm, i, 0), vp, fmx_, fmy_);
    }
    out[c * 2 + 0] = fmx_;
    out[c * 2 + 1] = fmy_;
  }
}

// face_velocity_gradients (fluxes.py:227-238): (dv/dx, dv/dy), each (ny, nx, 2), at the
// low face of each cell along `axis`: the normal difference of the two cells at the face
// and the average of their centred transverse differences, in the expression order of
// viscous_faces_kernel (so the gradients are the ones the viscous flux uses).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) face_gradients_kernel(const T* W, T* dvdx, T* dvdy,
                                                                   int nx, int ny, int nv,
                                                                   int axis,
                                                                   Over<T, false> over_dx,
                                                                   Over<T, false> over_dy,
                                                                   Over<T, false> over_4dx,
                                                                   Over<T, false> over_4dy) {
  const size_t n = (size_t)nx * ny;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c / nx), i = (int)(c % nx);
    const int im = (i == 0) ? nx - 1 : i - 1, ip = (i == nx - 1) ? 0 : i + 1;
    const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j == ny - 1) ? 0 : j + 1;
    auto w = [&](int jj, int ii, int v) { return W[((size_t)jj * nx + ii) * nv + v]; };
    for (int v = 1; v <= 2; ++v) {
      if (axis == 1) {
        dvdx[c * 2 + v - 1] = over_dx(w(j, i, v) - w(j, im, v));
        dvdy[c * 2 + v - 1] = over_4dy((w(jp, i, v) + w(jp, im, v)) - (w(jm, i, v) + w(jm, im, v)));
      } else {
        dvdy[c * 2 + v - 1] = over_dy(w(j, i, v) - w(jm, i, v));
        dvdx[c * 2 + v - 1] = over_4dx((w(j, ip, v) + w(jm, ip, v)) - (w(j, im, v) + w(jm, im, v)));
      }
    }
  }
}

// smagorinsky_viscosity (fluxes.py:241-245) from face gradients: c sqrt(2 (ux^2 + vy^2) +
// (uy + vx)^2) with c = Cs^2 dx dy rounded once on the host, as smag_nu in the fused stage.
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) smagorinsky_kernel(const T* dvdx, const T* dvdy,
                                                                T* nu, size_t n, T c) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    nu[k] = smag_nu(dvdx[2 * k], dvdy[2 * k + 1], dvdx[2 * k + 1] + dvdy[2 * k], c);
}

// flux[..., 1:3] += add (the viscous and learned-stress additions, hydro.py:106-119).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) add_momentum_kernel(T* flux, const T* add,
                                                                 size_t n, int nv) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    flux[c * nv + 1] = flux[c * nv + 1] + add[c * 2];
    flux[c * nv + 2] = flux[c * nv + 2] + add[c * 2 + 1];
  }
}

// Drag with the stage's incoming primitives, U[..., 1:3] -= (beta dt rate rho) v
// (hydro.py:122-125); beta dt is rounded once, as in the fused stage.
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) drag_kernel(T* U, const T* W, size_t n, int nv,
                                                         T coef) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T k = coef * W[c * nv];
    U[c * nv + 1] = U[c * nv + 1] + (-(k * W[c * nv + 1]));
    U[c * nv + 2] = U[c * nv + 2] + (-(k * W[c * nv + 2]));
  }
}

// The kick U[..., 1:3] += rho * (acceleration * dt) (hydro.py:131-132).
template <typename T>
__global__ void __launch_bounds__(kAuxBlock) kick_add_kernel(T* U, const T* acc, size_t n,
                                                             int nv, T dt) {
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const T rho = U[c * nv];
    U[c * nv + 1] = U[c * nv + 1] + rho * (acc[c * 2] * dt);
    U[c * nv + 2] = U[c * nv + 2] + rho * (acc[c * 2 + 1] * dt);
  }
}

}  // namespace jx

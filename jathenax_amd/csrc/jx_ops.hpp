// Scalar face operators of the isothermal update, shared by every gfx950 kernel.
//
// Each function restates one reference operator (jathenax/fluxes.py) for one cell or
// face, in the reference's floating-point expression order.  The float64 translation
// units are built with -ffp-contract=off, so in FP64 these are the IEEE evaluations of
// the reference expressions.  In FP32 (Fast<float>) divisions by run constants become
// reciprocal multiplies and sqrt/rcp use the hardware approximations, and the FP32 stage
// units fuse a product only into an add of the same expression (-ffp-contract=on), so an
// FP32 result does not depend on the tile shape; FP32 results are checked against the FP64
// form by error norms, not bit patterns.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "jx_types.hpp"

namespace jx {

// Pointers read from the instance table are generic (flat) to the compiler.  Flat accesses
// also count against the LDS wait counter, so every LDS wait in the row march would drain
// the row prefetch; casting them to the global address space keeps them on the vector
// memory path.
template <typename T> using gmem_t = T __attribute__((address_space(1)));
template <typename T> __device__ __forceinline__ gmem_t<T>* gmem(T* p) {
  return (gmem_t<T>*)p;
}
template <typename T> __device__ __forceinline__ const gmem_t<T>* gmem(const T* p) {
  return (const gmem_t<T>*)p;
}

// Row access through a buffer resource.  The resource is built per row from the row's
// wave-uniform base (scalar instructions only) and each lane adds a 32-bit byte offset, so
// a row load or store costs no vector address arithmetic; a plain pointer access makes the
// compiler fold the loop-invariant lane offset into a 64-bit vector base and add the row
// offset to it with a 64-bit vector add per access.  num_records is the row's byte length
// (an access past it reads zeros and writes nothing); dword 3 is the GFX9 raw-buffer
// format word (32-bit data format).
constexpr int kRawBufferWord3 = 0x00020000;

template <typename T> __device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const T* row,
                                                                              int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(row), (short)0, bytes,
                                           kRawBufferWord3);
}

// NV values of type T at byte offset off (NV * sizeof(T) bytes, 4-byte aligned), loaded
// and stored in chunks of at most 16 bytes.
template <typename T, int NV, int AUX = 0>
__device__ __forceinline__ void row_load(__amdgpu_buffer_rsrc_t r, unsigned off, T (&out)[NV]) {
  constexpr int W = NV * (int)sizeof(T) / 4;
  unsigned d[W];
  int k = 0;
#pragma unroll
  for (; k + 4 <= W; k += 4) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(off + 4 * k), 0, AUX);
    d[k] = x[0]; d[k + 1] = x[1]; d[k + 2] = x[2]; d[k + 3] = x[3];
  }
  if constexpr (W % 4 == 3) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b96(r, (int)(off + 4 * k), 0, AUX);
    d[k] = x[0]; d[k + 1] = x[1]; d[k + 2] = x[2];
  } else if constexpr (W % 4 == 2) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 4 * k), 0, AUX);
    d[k] = x[0]; d[k + 1] = x[1];
  } else if constexpr (W % 4 == 1) {
    d[k] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 4 * k), 0, AUX);
  }
  __builtin_memcpy(out, d, sizeof(d));
}

// Cache-policy word of a non-temporal (streaming) buffer access on gfx950: the line is
// not kept in L2 for reuse.  The stage kernels store their output rows this way (no later
// access of the kernel reads them): 8 % faster per launch at 4096^2 FP32 than plain stores.
// Their loads stay plain: the input rows are read again by neighbouring strips and bands,
// and non-temporal loads measured slower for both the input and the saved register.
constexpr int kNonTemporal = 2;

template <typename T, int NV, int AUX = 0>
__device__ __forceinline__ void row_store(__amdgpu_buffer_rsrc_t r, unsigned off,
                                          const T (&in)[NV]) {
  constexpr int W = NV * (int)sizeof(T) / 4;
  unsigned d[W];
  __builtin_memcpy(d, in, sizeof(d));
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  typedef unsigned u3 __attribute__((ext_vector_type(3)));
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  int k = 0;
#pragma unroll
  for (; k + 4 <= W; k += 4) {
    const u4 x = {d[k], d[k + 1], d[k + 2], d[k + 3]};
    __builtin_amdgcn_raw_buffer_store_b128(x, r, (int)(off + 4 * k), 0, AUX);
  }
  if constexpr (W % 4 == 3) {
    const u3 x = {d[k], d[k + 1], d[k + 2]};
    __builtin_amdgcn_raw_buffer_store_b96(x, r, (int)(off + 4 * k), 0, AUX);
  } else if constexpr (W % 4 == 2) {
    const u2 x = {d[k], d[k + 1]};
    __builtin_amdgcn_raw_buffer_store_b64(x, r, (int)(off + 4 * k), 0, AUX);
  } else if constexpr (W % 4 == 1) {
    __builtin_amdgcn_raw_buffer_store_b32(d[k], r, (int)(off + 4 * k), 0, AUX);
  }
}

// The instance table (written by the host between rounds) and the work lists (written by
// the control kernel, an earlier launch) are read-only while a worker kernel runs, so they
// are read through the constant address space: uniform addresses then become scalar loads
// into SGPRs instead of vector loads into VGPRs.
template <typename T> using cmem_t = T __attribute__((address_space(4)));
template <typename S> __device__ __forceinline__ S load_const(const S* p) {
  static_assert(sizeof(S) % 4 == 0, "record must be a whole number of dwords");
  S out;
  const cmem_t<int>* src = (const cmem_t<int>*)p;
  int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(S) / 4); ++i) dst[i] = src[i];
  return out;
}

template <typename T> struct Fast { static constexpr bool value = false; };
template <> struct Fast<float> { static constexpr bool value = true; };

// a / b: IEEE division in FP64, v_rcp_f32-based in FP32.
template <typename T> __device__ __forceinline__ T dvd(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float dvd<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
// x / c for a compile-time constant c (the reference's "/ 12.0", "/ 6.0").
template <typename T> __device__ __forceinline__ T divk(T x, double c) {
  if constexpr (Fast<T>::value) return x * T(1.0 / c);
  else return x / T(c);
}
template <typename T> __device__ __forceinline__ T sqr(T x) { return x * x; }
template <typename T> __device__ __forceinline__ T root(T x) { return sqrt(x); }
template <> __device__ __forceinline__ float root<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
template <typename T> __device__ __forceinline__ T fmn(T a, T b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T fmx(T a, T b) { return fmax(a, b); }
// fmin(fmax(e, fmin(a, b)), fmax(a, b)): clamp e between a and b (one v_med3 in FP32).
template <typename T> __device__ __forceinline__ T clamp_between(T e, T a, T b) {
  if constexpr (Fast<T>::value) return __builtin_amdgcn_fmed3f(e, a, b);
  else return fmn(fmx(e, fmn(a, b)), fmx(a, b));
}

// Density floor (x < f) ? f : x (hydro.py:56). FP32 uses the NaN-propagating maximum, one
// instruction with the same result for every x, NaN included (f > 0).
template <typename T> __device__ __forceinline__ T floor_to(T x, T f) {
  if constexpr (Fast<T>::value) return __builtin_elementwise_maximum(x, f);
  else return (x < f) ? f : x;
}

// Division by a run constant d (dx, dy, cs, 4dx, 4dy).  EXACT: d is a power of two, so
// x / d == x * (1/d) bit for bit and the multiply is used even in FP64.
template <typename T, bool EXACT> struct Over {
  T d, inv;
  __device__ __forceinline__ T operator()(T x) const {
    if constexpr (Fast<T>::value || EXACT) return x * inv;
    else return x / d;
  }
};

// PPM4 interface value between cells c-1 and c, clamped between them (fluxes.py:29-32).
// The reference evaluates it twice, as the low edge of c and the high edge of c-1; the
// operand order of both evaluations is the same up to commuted additions, so one value
// serves both cells.
template <typename T>
__device__ __forceinline__ T ppm4_interface(T qm2, T qm1, T q, T qp1) {
  return clamp_between(divk(T(7) * (q + qm1) - (qm2 + qp1), 12.0), q, qm1);
}

// PPM4 extremum and asymmetric-parabola limiter on clamped edges (fluxes.py:34-41).
// FP32 uses the same limiter as two medians: with dl = lo - q and dh = hi - q,
// lo' = med3(lo, q - 2 dh, q). At an extremum (dl, dh of one sign, or one zero) lo and
// q - 2 dh lie on either side of q and the median is q. Otherwise both lie on the side
// of lo and the median is the nearer to q: q - 2 dh exactly when |dl| >= 2 |dh|.
// Likewise for hi'. Each result is one of the select form's operands (lo, q - 2 dh or q).
template <typename T>
__device__ __forceinline__ void ppm4_limit(T q, T& lo, T& hi) {
  const T dh = hi - q, dl = lo - q;
  if constexpr (Fast<T>::value) {
    lo = __builtin_amdgcn_fmed3f(lo, q - T(2) * dh, q);
    hi = __builtin_amdgcn_fmed3f(hi, q - T(2) * dl, q);
    return;
  }
  const bool ext = dh * dl >= T(0);
  const T l2 = (fabs(dl) >= T(2) * fabs(dh)) ? q - T(2) * dh : lo;
  const T h2 = (fabs(dh) >= T(2) * fabs(dl)) ? q - T(2) * dl : hi;
  lo = ext ? q : l2;
  hi = ext ? q : h2;
}

// ---------------------------------------------------------------------------------------
// Reconstruction: low and high edge values of one cell from its 5-point stencil.  The
// reference's left state at face i is the high edge of cell i-1 and its right state the
// low edge of cell i (fluxes.py:6-117).
// ---------------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T ppmx_limit_edge(T a, T b, T e, T ca, T cb) {
  // fluxes.py:45-53
  const T curv = T(3) * ((a + b) - T(2) * e);
  const bool same = ((curv > T(0)) && (ca > T(0)) && (cb > T(0))) ||
                    ((curv < T(0)) && (ca < T(0)) && (cb < T(0)));
  const T mag = fmn(fabs(ca), fabs(cb));
  const T lim = copysign(fmn(T(1.25) * mag, fabs(curv)), curv);
  const T corr = T(0.5) * (a + b) - divk(same ? lim : T(0), 6.0);
  return ((a - e) * (b - e) > T(0)) ? corr : e;
}

// FP32 WENO-Z (fluxes.py:87-117) from first differences.  With dm2 = qm1 - qm2, dm = q - qm1,
// dp = qp1 - q, dp2 = qp2 - qp1 the smoothness indicators are
//   4 b_k = (13/3) c_k^2 + s_k^2,  c = (dm - dm2, dp - dm, dp2 - dp),
//                                  s = (3 dm - dm2, dm + dp, dp2 - 3 dp),
// and the weights depend only on ratios of the b_k + eps, so 4 b_k + 4 eps are used as they
// are.  The three candidate values at the high edge differ from the central one,
// p1 = q + (dm + 2 dp) / 6, by third differences: p0 - p1 = ta / 3 and p2 - p1 = tb / 6 with
// ta = c0 - c1, tb = c1 - c2 (at the low edge, p1' = q - (2 dm + dp) / 6, p0' - p1' = -tb / 3
// and p2' - p1' = -ta / 6).  With the linear weights (0.1, 0.6, 0.3) folded in:
//   hi = p1  + (w0 ta / 3 + w2 tb / 2) / (w0 + 6 w1 + 3 w2),
//   lo = p1' - (w2 tb / 3 + w0 ta / 2) / (w2 + 6 w1 + 3 w0).
// Same weights and edge values as the reference expressions, in about two thirds of the
// operations, and without their cancellation between large stencil sums: the float32
// error against the float64 reference form is about half that of the reference form in
// float32.  The weight sums are at least 1 (the smallest d_k gives w = b^2 + 1 when
// tau >= dmin, and base = 1 otherwise), so both reciprocals are finite.
__device__ __forceinline__ void wenoz_edges_fast(float qm2, float qm1, float q, float qp1,
                                                 float qp2, float eps, float& lo, float& hi) {
  const float dm2 = qm1 - qm2, dm = q - qm1, dp = qp1 - q, dp2 = qp2 - qp1;
  const float c0 = dm - dm2, c1 = dp - dm, c2 = dp2 - dp;
  const float s0 = fmaf(3.0f, dm, -dm2), s1 = dm + dp, s2 = fmaf(-3.0f, dp, dp2);
  const float k = 13.0f / 3.0f;
  const float b0 = fmaf(k * c0, c0, s0 * s0);
  const float b1 = fmaf(k * c1, c1, s1 * s1);
  const float b2 = fmaf(k * c2, c2, s2 * s2);
  const float e4 = 4.0f * eps;
  const float tau = fabsf(b0 - b2);
  const float d0 = b0 + e4, d1 = b1 + e4, d2 = b2 + e4;
  const float dmin = fminf(d0, fminf(d1, d2));
  const float b = dmin * __builtin_amdgcn_rcpf(fmaxf(tau, dmin));   // dmin / scale
  const float g = tau * b;                                           // tau dmin / scale
  const float base = b * b;
  const float t0 = g * __builtin_amdgcn_rcpf(d0), t1 = g * __builtin_amdgcn_rcpf(d1),
              t2 = g * __builtin_amdgcn_rcpf(d2);
  const float w0 = fmaf(t0, t0, base), w1 = fmaf(t1, t1, base), w2 = fmaf(t2, t2, base);
  const float ta = c0 - c1, tb = c1 - c2;
  const float u = w0 * ta, v = w2 * tb;
  hi = fmaf(fmaf(u, 1.0f / 3.0f, 0.5f * v),
            __builtin_amdgcn_rcpf(fmaf(3.0f, w2, fmaf(6.0f, w1, w0))),
            fmaf(1.0f / 6.0f, fmaf(2.0f, dp, dm), q));
  lo = fmaf(-fmaf(v, 1.0f / 3.0f, 0.5f * u),
            __builtin_amdgcn_rcpf(fmaf(3.0f, w0, fmaf(6.0f, w1, w2))),
            fmaf(-1.0f / 6.0f, fmaf(2.0f, dm, dp), q));
}

template <int RECON, typename T>
__device__ __forceinline__ void cell_edges(T qm2, T qm1, T q, T qp1, T qp2, T eps, T& lo,
                                           T& hi) {
  if constexpr (RECON == RECON_DC) {
    // fluxes.py:6-8
    lo = q;
    hi = q;
  } else if constexpr (RECON == RECON_PLM) {
    // fluxes.py:11-18
    const T dl = q - qm1, dr = qp1 - q;
    const T p = dl * dr;
    const T s = (p > T(0)) ? dvd(p, dl + dr) : T(0);
    hi = q + s;
    lo = q - s;
  } else if constexpr (RECON == RECON_PPM4) {
    // fluxes.py:21-42
    lo = ppm4_interface(qm2, qm1, q, qp1);
    hi = ppm4_interface(qm1, q, qp1, qp2);
    ppm4_limit(q, lo, hi);
  } else if constexpr (RECON == RECON_PPMX) {
    // fluxes.py:56-84
    T l = divk(T(7) * (q + qm1) - (qm2 + qp1), 12.0);
    T h = divk(T(7) * (q + qp1) - (qm1 + qp2), 12.0);
    const T d2l = (qm2 + q) - T(2) * qm1;
    const T d2c = (qm1 + qp1) - T(2) * q;
    const T d2r = (q + qp2) - T(2) * qp1;
    l = ppmx_limit_edge(qm1, q, l, d2l, d2c);
    h = ppmx_limit_edge(q, qp1, h, d2c, d2r);
    const T dh = h - q, dl = l - q;
    const bool ext = (dh * (q - l) <= T(0)) || ((qm1 - q) * (q - qp1) <= T(0));
    const T d2 = T(6) * ((l + h) - T(2) * q);
    const bool same = ((d2 > T(0)) && (d2l > T(0)) && (d2c > T(0)) && (d2r > T(0))) ||
                      ((d2 < T(0)) && (d2l < T(0)) && (d2c < T(0)) && (d2r < T(0)));
    const T mag = fmn(fabs(d2c), fmn(fabs(d2l), fabs(d2r)));
    const T lim = copysign(fmn(T(1.25) * mag, fabs(d2)), d2);
    const T scale = fmx(fabs(qm1), fmx(fabs(q), fabs(qp1)));
    const bool use = same && (fabs(d2) > T(1e-12) * scale);
    const T ratio = use ? dvd(lim, d2) : T(0);
    const T lm = (fabs(dl) >= T(2) * fabs(dh)) ? q - T(2) * dh : l;
    const T hm = (fabs(dh) >= T(2) * fabs(dl)) ? q - T(2) * dl : h;
    lo = ext ? q + dl * ratio : lm;
    hi = ext ? q + dh * ratio : hm;
  } else {
    // WENO-Z, fluxes.py:87-117
    if constexpr (Fast<T>::value) {
      wenoz_edges_fast(qm2, qm1, q, qp1, qp2, eps, lo, hi);
      return;
    }
    const T c13 = T(13.0 / 12.0);
    const T b0 = c13 * sqr((qm2 + q) - T(2) * qm1) + T(0.25) * sqr((qm2 + T(3) * q) - T(4) * qm1);
    const T b1 = c13 * sqr((qm1 + qp1) - T(2) * q) + T(0.25) * sqr(qm1 - qp1);
    const T b2 = c13 * sqr((qp2 + q) - T(2) * qp1) + T(0.25) * sqr((qp2 + T(3) * q) - T(4) * qp1);
    const T tau = fabs(b0 - b2);
    const T d0 = b0 + eps, d1 = b1 + eps, d2 = b2 + eps;
    const T dmin = fmn(d0, fmn(d1, d2));
    const T scale = fmx(tau, dmin);
    const T base = sqr(dvd(dmin, scale));
    const T contrast = dvd(tau, scale);
    const T w0 = base + sqr(contrast * dvd(dmin, d0));
    const T w1 = base + sqr(contrast * dvd(dmin, d1));
    const T w2 = base + sqr(contrast * dvd(dmin, d2));
    {
      const T a0 = T(0.1) * w0, a1 = T(0.6) * w1, a2 = T(0.3) * w2;
      const T f0 = (T(2) * qm2 - T(7) * qm1) + T(11) * q;
      const T f1 = (-qm1 + T(5) * q) + T(2) * qp1;
      const T f2 = (T(2) * q + T(5) * qp1) - qp2;
      hi = dvd((f0 * a0 + f1 * a1) + f2 * a2, T(6) * ((a0 + a1) + a2));
    }
    {
      const T a0 = T(0.1) * w2, a1 = T(0.6) * w1, a2 = T(0.3) * w0;
      const T f0 = (T(2) * qp2 - T(7) * qp1) + T(11) * q;
      const T f1 = (-qp1 + T(5) * q) + T(2) * qm1;
      const T f2 = (T(2) * q + T(5) * qm1) - qm2;
      lo = dvd((f0 * a0 + f1 * a1) + f2 * a2, T(6) * ((a0 + a1) + a2));
    }
  }
}

// PPMX and WENO-Z floor the face density (fluxes.py:128-130).
template <int RECON> __host__ __device__ constexpr bool floors_faces() {
  return RECON == RECON_PPMX || RECON == RECON_WENOZ;
}

// ---------------------------------------------------------------------------------------
// Isothermal Riemann solvers in (normal, tangent) order: returns mass, normal momentum and
// tangential momentum fluxes (fluxes.py:134-201).
// ---------------------------------------------------------------------------------------

template <typename T> struct Gas {
  T cs, cs2, inv_cs;
};

// Roe averages (sqrt(rl) a_l + sqrt(rr) a_r) / (sqrt(rl) + sqrt(rr)) (fluxes.py:145-148,
// 187-188).  FP64 evaluates the reference expression.  FP32 uses the equal weight
// sqrt(rr) / (sqrt(rl) + sqrt(rr)) = rr / (sqrt(rl rr) + rr) and a_l + (a_r - a_l) w: one
// square root and one reciprocal instead of two square roots and a reciprocal (both are
// quarter-rate instructions).
template <typename T>
__device__ __forceinline__ void roe_averages(T rl, T rr, T ul, T ur, T tl, T tr, T& vn, T& vt) {
  if constexpr (Fast<T>::value) {
    const T w = rr * dvd(T(1), root(rl * rr) + rr);
    vn = ul + (ur - ul) * w;
    vt = tl + (tr - tl) * w;
  } else {
    const T sl = root(rl), sr = root(rr);
    const T inv = dvd(T(1), sl + sr);
    vn = (sl * ul + sr * ur) * inv;
    vt = (sl * tl + sr * tr) * inv;
  }
}
template <typename T>
__device__ __forceinline__ T roe_average(T rl, T rr, T ul, T ur) {
  if constexpr (Fast<T>::value) {
    const T w = rr * dvd(T(1), root(rl * rr) + rr);
    return ul + (ur - ul) * w;
  } else {
    const T sl = root(rl), sr = root(rr);
    return (sl * ul + sr * ur) * dvd(T(1), sl + sr);
  }
}

// FP32 Roe flux (fluxes.py:134-172) in fewer operations; equal to the reference form in
// exact arithmetic:
//  * it forms twice the flux and halves it at the end, grouping the left and right
//    physical fluxes into one sum per component;
//  * no supersonic selects: when every wave speed has one sign, |lambda| = +-lambda and
//    the Roe property A(Ur - Ul) = Fr - Fl make the upwind sum equal fl (or fr) up to
//    rounding;
//  * the LLF fallback (negative intermediate or right density) is rare, so one wave-wide
//    vote skips it; the empty volatile asm keeps the compiler from turning the skipped
//    block back into selects that every face would execute.
// X2: return twice the flux (the caller folds the halving into its flux coefficient).
template <bool X2, typename T>
__device__ __forceinline__ void roe_fast(T rl, T rr, T ul, T ur, T tl, T tr, const Gas<T>& g,
                                         T& f0, T& f1, T& f2) {
  const T cs = g.cs;
  const T mnl = rl * ul, mnr = rr * ur, mtl = rl * tl;
  T vn, vt;
  roe_averages(rl, rr, ul, ur, tl, tr, vn, vt);
  const T wm = vn - cs, wp = vn + cs;
  const T drho = rr - rl, dmn = mnr - mnl, dmt = rr * tr - mtl;
  const T h = T(0.5) * g.inv_cs;
  const T am = drho * (T(0.5) + vn * h) - dmn * h;   // wave strengths, am + ap = drho
  const T ap = drho - am;
  const T ash = dmt - vt * drho;
  const T s0 = mnl + mnr;
  const T s1 = (mnl * ul + mnr * ur) + g.cs2 * (rl + rr);
  const T s2 = mnl * tl + mnr * tr;
  const T A = fabs(wm) * am, B = fabs(wp) * ap, AB = A + B;
  T m0 = s0 - AB;
  T m1 = (s1 - A * wm) - B * wp;
  T m2 = (s2 - AB * vt) - fabs(vn) * ash;
  const bool bad = (rl + am < T(0)) || (rr < T(0));
  if (__builtin_amdgcn_ballot_w64(bad) != 0) {
    __asm__ volatile("");
    if (bad) {
      const T sp = fmx(fabs(ul) + cs, fabs(ur) + cs);
      m0 = s0 - sp * drho;
      m1 = s1 - sp * dmn;
      m2 = s2 - sp * dmt;
    }
  }
  if constexpr (X2) {
    f0 = m0;
    f1 = m1;
    f2 = m2;
  } else {
    f0 = T(0.5) * m0;
    f1 = T(0.5) * m1;
    f2 = T(0.5) * m2;
  }
}

// The FP32 fused and batched stage kernels take the Roe and LLF fluxes at twice their
// value: the solvers' final halving is folded into the update's flux coefficient (and the
// viscous and learned-stress face terms are doubled to match, both exactly).  FP64 and
// HLLE return the flux itself.
template <int RS, typename T> constexpr bool flux_x2() {
  return Fast<T>::value && RS != RS_HLLE;
}

// Riemann flux at one face; X2 (only where flux_x2<RS, T>()): twice the flux.
template <int RS, bool EXACT, bool X2, typename T>
__device__ __forceinline__ void riemann_x(T rl, T rr, T ul, T ur, T tl, T tr, const Gas<T>& g,
                                          T& f0, T& f1, T& f2) {
  static_assert(!X2 || flux_x2<RS, T>(), "doubled flux only for FP32 Roe and LLF");
  const T cs = g.cs, cs2 = g.cs2;
  if constexpr (RS == RS_LLF) {
    // fluxes.py:175-183
    const T mnl = rl * ul, mnr = rr * ur;
    const T mtl = rl * tl, mtr = rr * tr;
    const T s0 = mnl + mnr;
    const T s1 = (mnl * ul + mnr * ur) + cs2 * (rl + rr);
    const T s2 = mnl * tl + mnr * tr;
    const T sp = fmx(fabs(ul) + cs, fabs(ur) + cs);
    if constexpr (X2) {
      f0 = s0 - sp * (rr - rl);
      f1 = s1 - sp * (mnr - mnl);
      f2 = s2 - sp * (mtr - mtl);
      return;
    }
    f0 = T(0.5) * (s0 - sp * (rr - rl));
    f1 = T(0.5) * (s1 - sp * (mnr - mnl));
    f2 = T(0.5) * (s2 - sp * (mtr - mtl));
  } else if constexpr (RS == RS_HLLE) {
    // fluxes.py:186-201
    const T vn = roe_average(rl, rr, ul, ur);
    if constexpr (Fast<T>::value) {
      // FP32: min(a - c, b - c) = min(a, b) - c exactly (rounding is monotone), and
      // 0.5 (fl + fr) + wt (fl - fr) = (bp fl - bm fr) / (bp - bm) with the reference's
      // shifted fluxes fl = F(Ul) - bm Ul, fr = F(Ur) - bp Ur, whose momentum and
      // tangential parts are fl0 ul + cs^2 rl and fl0 tl.
      const T al = fmn(vn, ul) - cs;
      const T ar = fmx(vn, ur) + cs;
      const T bm = (al < T(0)) ? al : T(-1e-20);
      const T bp = (ar > T(0)) ? ar : T(1e-20);
      const T fl0 = rl * (ul - bm), fr0 = rr * (ur - bp);
      const T fl1 = fl0 * ul + cs2 * rl, fr1 = fr0 * ur + cs2 * rr;
      const T fl2 = fl0 * tl, fr2 = fr0 * tr;
      const T inv = dvd(T(1), bp - bm);
      const T cl = bp * inv, cr = -(bm * inv);
      f0 = cl * fl0 + cr * fr0;
      f1 = cl * fl1 + cr * fr1;
      f2 = cl * fl2 + cr * fr2;
      return;
    }
    const T al = fmn(vn - cs, ul - cs);
    const T ar = fmx(vn + cs, ur + cs);
    const T bm = (al < T(0)) ? al : T(-1e-20);
    const T bp = (ar > T(0)) ? ar : T(1e-20);
    const T dl = ul - bm, dr = ur - bp;
    const T fl0 = rl * dl, fl1 = (rl * ul) * dl + cs2 * rl, fl2 = (rl * tl) * dl;
    const T fr0 = rr * dr, fr1 = (rr * ur) * dr + cs2 * rr, fr2 = (rr * tr) * dr;
    const T wt = dvd(T(0.5) * (bp + bm), bp - bm);
    f0 = T(0.5) * (fl0 + fr0) + wt * (fl0 - fr0);
    f1 = T(0.5) * (fl1 + fr1) + wt * (fl1 - fr1);
    f2 = T(0.5) * (fl2 + fr2) + wt * (fl2 - fr2);
  } else {
    // Roe with supersonic upwinding and the LLF fallback, fluxes.py:134-172
    if constexpr (Fast<T>::value) {
      roe_fast<X2>(rl, rr, ul, ur, tl, tr, g, f0, f1, f2);
      return;
    }
    const T mnl = rl * ul, mnr = rr * ur;
    const T mtl = rl * tl, mtr = rr * tr;
    const T fl0 = mnl, fl1 = mnl * ul + cs2 * rl, fl2 = mnl * tl;
    const T fr0 = mnr, fr1 = mnr * ur + cs2 * rr, fr2 = mnr * tr;
    const T fa0 = T(0.5) * (fl0 + fr0), fa1 = T(0.5) * (fl1 + fr1), fa2 = T(0.5) * (fl2 + fr2);
    T vn, vt;
    roe_averages(rl, rr, ul, ur, tl, tr, vn, vt);
    const T wm = vn - cs, wp = vn + cs;
    const T drho = rr - rl, dmn = mnr - mnl, dmt = mtr - mtl;
    T am, ap;
    if constexpr (Fast<T>::value) {
      // am + ap = drho exactly, so FP32 forms ap from am; 0.5/cs is folded into one factor.
      const T h = T(0.5) * g.inv_cs;
      am = drho * (T(0.5) + vn * h) - dmn * h;
      ap = drho - am;
    } else if constexpr (EXACT) {
      am = drho * (T(0.5) + T(0.5) * vn * g.inv_cs) - dmn * T(0.5) * g.inv_cs;
      ap = drho * (T(0.5) - T(0.5) * vn * g.inv_cs) + dmn * T(0.5) * g.inv_cs;
    } else {
      am = drho * (T(0.5) + (T(0.5) * vn) / cs) - (dmn * T(0.5)) / cs;
      ap = drho * (T(0.5) - (T(0.5) * vn) / cs) + (dmn * T(0.5)) / cs;
    }
    const T ash = drho * (-vt) + dmt;
    const T cm = (T(-0.5) * fabs(wm)) * am;
    const T csh = (T(-0.5) * fabs(vn)) * ash;
    const T cp = (T(-0.5) * fabs(wp)) * ap;
    T m0 = (fa0 + cm) + cp;
    T m1 = (fa1 + cm * wm) + cp * wp;
    T m2 = ((fa2 + cm * vt) + csh) + cp * vt;
    if (wm >= T(0)) { m0 = fl0; m1 = fl1; m2 = fl2; }
    if (wp <= T(0)) { m0 = fr0; m1 = fr1; m2 = fr2; }
    const T rstar = rl + am;
    // rstar + ap = rl + drho = rr; FP32 tests rr itself.
    const bool neg2 = Fast<T>::value ? (rr < T(0)) : (rstar + ap < T(0));
    if ((rstar < T(0)) || neg2) {
      const T sp = T(0.5) * fmx(fabs(ul) + cs, fabs(ur) + cs);
      m0 = fa0 - sp * drho;
      m1 = fa1 - sp * dmn;
      m2 = fa2 - sp * dmt;
    }
    f0 = m0;
    f1 = m1;
    f2 = m2;
  }
}

// The flux itself (fluxes.py:204-219 per face).
template <int RS, bool EXACT, typename T>
__device__ __forceinline__ void riemann(T rl, T rr, T ul, T ur, T tl, T tr, const Gas<T>& g,
                                        T& f0, T& f1, T& f2) {
  riemann_x<RS, EXACT, false>(rl, rr, ul, ur, tl, tr, g, f0, f1, f2);
}

// ---------------------------------------------------------------------------------------
// Viscous momentum flux at one face (fluxes.py:227-266).  Inputs are the face velocity
// gradients dvx/dx, dvy/dy, dvx/dy, dvy/dx and rho_i + rho_(i-1).  Returns the x and y
// momentum flux increments -mu * stress.  AXIS follows the reference: 1 = x faces.
// ---------------------------------------------------------------------------------------

template <typename T> struct Visc {
  T nu;       // uniform kinematic viscosity
  T smag_c;   // Cs^2 dx dy (0: Smagorinsky off)
  bool smag;
};

template <typename T>
__device__ __forceinline__ T smag_nu(T ux, T vy, T shear, T c) {
  // fluxes.py:241-245
  const T s2 = T(2) * (ux * ux + vy * vy) + shear * shear;
  return c * root(s2);
}

// h scales mu = (h nu) (rho_i + rho_(i-1)): 0.5 for the reference's flux, 1 for twice it
// (exactly, as the scaling is by a power of two).
template <int AXIS, typename T>
__device__ __forceinline__ void visc_face_h(T ux, T vy, T uy, T vx, T rho_sum, const Visc<T>& p,
                                            T h, T& fmx_out, T& fmy_out) {
  const T shear = vx + uy;    // dv_dx[..., 1] + dv_dy[..., 0]
  T sx, sy;
  if constexpr (AXIS == 1) {
    sx = T(4.0 / 3.0) * ux - T(2.0 / 3.0) * vy;
    sy = shear;
  } else {
    sx = shear;
    sy = T(4.0 / 3.0) * vy - T(2.0 / 3.0) * ux;
  }
  T nu = p.nu;
  if (p.smag) nu = p.nu + smag_nu(ux, vy, shear, p.smag_c);
  const T mu = (h * nu) * rho_sum;
  fmx_out = (-mu) * sx;
  fmy_out = (-mu) * sy;
}

// X2: twice the increments (h = 1; the compiler drops the multiply by one).
template <int AXIS, bool X2 = false, typename T>
__device__ __forceinline__ void visc_face(T ux, T vy, T uy, T vx, T rho_sum, const Visc<T>& p,
                                          T& fmx_out, T& fmy_out) {
  visc_face_h<AXIS>(ux, vy, uy, vx, rho_sum, p, X2 ? T(1) : T(0.5), fmx_out, fmy_out);
}

// ---------------------------------------------------------------------------------------
// FP32 conservative update of one zone (hydro.py:39-51 with fluxes.py:283-287), shared by
// the fused and the batched stage kernels so that both round alike:
//   u' = (base - cx (F_(i+1) - F_i)) - cy (G_(j+1) - G_j),
// with cx = beta dt / dx and cy = beta dt / dy (halved when the fluxes are doubled) and
// base the RK row's combination of the stage input u and the saved register us: u itself
// in stage 1 (every first row is (0, 1) and us == u there), us + g0 (u - us) for the
// blended rows, g0 u + g1 us otherwise.  The uniform branches sit outside the per-variable
// arithmetic.
// ---------------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T flux_coef(double beta_dt, T inv_d, bool x2) {
  return T(beta_dt * (double)inv_d * (x2 ? 0.5 : 1.0));
}

template <int NV, typename T>
__device__ __forceinline__ void rk_update_fast(const T (&u)[NV], const T (&us)[NV],
                                               const T (&dF)[NV], const T (&dG)[NV],
                                               bool stage1, bool blended, T g0, T g1, T cx,
                                               T cy, T (&out)[NV]) {
  T base[NV];
  if (stage1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) base[v] = u[v];
  } else if (blended) {
#pragma unroll
    for (int v = 0; v < NV; ++v) base[v] = us[v] + g0 * (u[v] - us[v]);
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) base[v] = g0 * u[v] + g1 * us[v];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const T t = base[v] - cx * dF[v];
    out[v] = t - cy * dG[v];
  }
}

// ---------------------------------------------------------------------------------------
// Order-preserving keys for the grid-wide dt min / nu max (NaN maps to 0 so that it wins
// a min, mirroring jnp.min's NaN propagation).
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long key_of(double x) {
  if (x != x) return 0ull;
  return static_cast<unsigned long long>(__double_as_longlong(x)) + 1ull;
}
__host__ __device__ inline double value_of_key(unsigned long long k) {
  if (k == 0ull) return __builtin_nan("");
  union { unsigned long long u; double d; } c;
  c.u = k - 1ull;
  return c.d;
}

// Per-lane CFL limit (hydro.py:66-68): min over the lane's zones of dx/(|vx| + cs) and
// dy/(|vy| + cs), handed on as a key. FP64 keys every zone's length as the reference
// divides. FP32 keeps the largest signal rate (|v| + cs) / d with a NaN-propagating
// maximum (one v_maximum3 per zone, no divisions) and keys its reciprocal, in float64,
// once per lane; a NaN state still yields key 0. Every FP32 kernel that forms the CFL
// limit uses this accumulator, so all of them agree bit for bit.
template <typename T> struct CflMin {
  T dx, dy, cs;
  unsigned long long k = ~0ull;
  __device__ CflMin(T dx_, T dy_, T cs_) : dx(dx_), dy(dy_), cs(cs_) {}
  __device__ __forceinline__ void add(T vx, T vy) {
    const unsigned long long kx = key_of((double)dvd(dx, fabs(vx) + cs));
    const unsigned long long ky = key_of((double)dvd(dy, fabs(vy) + cs));
    k = kx < k ? kx : k;
    k = ky < k ? ky : k;
  }
  __device__ __forceinline__ unsigned long long key() const { return k; }
};
template <> struct CflMin<float> {
  float inv_dx, inv_dy, cs;
  float smax = 0.0f;
  __device__ CflMin(float dx_, float dy_, float cs_)
      : inv_dx((float)(1.0 / (double)dx_)), inv_dy((float)(1.0 / (double)dy_)), cs(cs_) {}
  __device__ __forceinline__ void add(float vx, float vy) {
    const float rx = (fabsf(vx) + cs) * inv_dx, ry = (fabsf(vy) + cs) * inv_dy;
    smax = __builtin_elementwise_maximum(__builtin_elementwise_maximum(smax, rx), ry);
  }
  __device__ __forceinline__ unsigned long long key() const {
    // Rates are positive, so 0 means the lane met no zone.
    return smax == 0.0f ? ~0ull : key_of(1.0 / (double)smax);
  }
};

// ---------------------------------------------------------------------------------------
// FP32 raw force evaluated where it is used (forcing.py:111-122).  The grouped render
// writes, per component, f(j, i) = sum over groups g of cx_g(i) P_g(j) - sx_g(i) Q_g(j), for
// groups of modes that share their x basis rows (up to the sign of sx).  With at most
// kMaxForceGroups groups only the per-row sums (P0, Q0, P1, Q1) are stored
// (force_rows_kernel, (ny, kMaxForceGroups, 4) values), and a kernel that needs the force
// evaluates it per cell from its column's x rows in the render's operation order: the same
// values, bit for bit, without the (ny, nx, 2) array.
// ---------------------------------------------------------------------------------------

constexpr int kMaxForceGroups = 4;

template <typename T> struct ForceEval {
  const T* pq;                 // per-row group sums; nullptr: the kernel reads a raw array
  const T* cx; const T* sx;    // basis x rows (M, nx)
  const int* grp;              // mode groups: G, gstart[M+1], order[M], ...
  int n_modes;
};

// Column i's x rows of each group's first mode (zero past the group count G).
template <typename T> struct ForceCols {
  T cx[kMaxForceGroups], sx[kMaxForceGroups];
  int G;
};

// The group table and the row sums are not written while a kernel reads them, and every
// lane of a wave reads the same words: they are read through the constant address space
// (scalar loads).
template <typename T>
__device__ __forceinline__ ForceCols<T> force_cols(const ForceEval<T>& f, int nx, int i) {
  ForceCols<T> c;
  const cmem_t<int>* grp = (const cmem_t<int>*)f.grp;
  c.G = grp[0];
  const cmem_t<int>* gstart = grp + 1;
  const cmem_t<int>* order = gstart + (f.n_modes + 1);
#pragma unroll
  for (int g = 0; g < kMaxForceGroups; ++g) {
    c.cx[g] = c.sx[g] = T(0);
    if (g < c.G) {
      const size_t m = (size_t)order[gstart[g]];
      c.cx[g] = f.cx[m * nx + i];
      c.sx[g] = f.sx[m * nx + i];
    }
  }
  return c;
}

template <typename T> __device__ __forceinline__ T fma_of(T a, T b, T c) {
  if constexpr (std::is_same<T, float>::value) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// The raw force of row j at the column of `c`: render_grouped_kernel's accumulation.
template <typename T>
__device__ __forceinline__ void force_at(const ForceEval<T>& f, const ForceCols<T>& c, int j,
                                         T& f0, T& f1) {
  const cmem_t<T>* row = (const cmem_t<T>*)f.pq + (size_t)j * (kMaxForceGroups * 4);
  f0 = f1 = T(0);
#pragma unroll
  for (int g = 0; g < kMaxForceGroups; ++g) {
    if (g < c.G) {
      f0 = fma_of(c.cx[g], row[4 * g + 0], fma_of(-c.sx[g], row[4 * g + 1], f0));
      f1 = fma_of(c.cx[g], row[4 * g + 2], fma_of(-c.sx[g], row[4 * g + 3], f1));
    }
  }
}

}  // namespace jx

// Host seed path of the turbulence driver, float64 (jathenax/forcing.py:11-100).
//
// ran_st is L'Ecuyer's two-stream combined generator with a 32-entry Bays-Durham shuffle
// (the Numerical Recipes "ran2" constants); ran_gaussian_st is the polar Box-Muller
// transform that returns the second value of each accepted pair first.  The modal table
// follows the reference's evaluation order; its parabolic spectrum uses a correctly
// rounded fma(prefactor, offset^2, 1), which is what the reference's exact Fraction
// arithmetic computes (forcing.py:76-82).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/jathenax_amd.h"

namespace {

class RanSt {
 public:
  explicit RanSt(long long seed) {
    s1_ = s2_ = seed > 1 ? seed : 1;
    for (int idx = 39; idx >= 0; --idx) {
      s1_ = (40014 * s1_) % kM1;
      if (idx < 32) table_[idx] = s1_;
    }
    iy_ = table_[0];
  }
  double next() {
    s1_ = (40014 * s1_) % kM1;
    s2_ = (40692 * s2_) % kM2;
    const int64_t j = iy_ / kDiv;
    iy_ = table_[j] - s2_;
    table_[j] = s1_;
    if (iy_ < 1) iy_ += kM1 - 1;
    const double v = (1.0 / (double)kM1) * (double)iy_;
    const double top = 1.0 - 2.220446049250313e-16;
    return v < top ? v : top;
  }

 private:
  static constexpr int64_t kM1 = 2147483563, kM2 = 2147483399;
  static constexpr int64_t kDiv = 1 + (kM1 - 1) / 32;
  int64_t s1_, s2_, iy_;
  int64_t table_[32];
};

class Gaussian {
 public:
  explicit Gaussian(long long seed) : u_(seed) {}
  double next() {
    if (have_) {
      have_ = false;
      return second_;
    }
    for (;;) {
      const double v1 = 2.0 * u_.next() - 1.0;
      const double v2 = 2.0 * u_.next() - 1.0;
      const double r2 = v1 * v1 + v2 * v2;
      if (0.0 < r2 && r2 < 1.0) {
        const double f = std::sqrt(-2.0 * std::log(r2) / r2);
        second_ = v1 * f;
        have_ = true;
        return v2 * f;
      }
    }
  }

 private:
  RanSt u_;
  bool have_ = false;
  double second_ = 0.0;
};

}  // namespace

extern "C" {

int jx_ran_st(long long seed, long long count, double* out) {
  if (count < 0 || (count > 0 && !out)) return JX_EINVAL;
  RanSt r(seed);
  for (long long i = 0; i < count; ++i) out[i] = r.next();
  return JX_OK;
}

int jx_ran_gaussian_st(long long seed, long long count, double* out) {
  if (count < 0 || (count > 0 && !out)) return JX_EINVAL;
  Gaussian g(seed);
  for (long long i = 0; i < count; ++i) out[i] = g.next();
  return JX_OK;
}

int jx_cartesian_modes(int nlow, int nhigh, int kx_min, int kx_max, int ky_min, int ky_max,
                       int* modes_out, int cap) {
  if (!(1 <= nlow && nlow < nhigh)) return -JX_EINVAL;
  int n = 0;
  for (int kx = kx_min; kx <= kx_max; ++kx)
    for (int ky = ky_min; ky <= ky_max; ++ky) {
      const long long k2 = (long long)kx * kx + (long long)ky * ky;
      if ((long long)nlow * nlow <= k2 && k2 <= (long long)nhigh * nhigh) {
        if (modes_out && n < cap) {
          modes_out[2 * n] = kx;
          modes_out[2 * n + 1] = ky;
        }
        ++n;
      }
    }
  if (n == 0) return -JX_EINVAL;
  return n;
}

int jx_modal_coefficients(const int* modes, int n_modes, double lx, double ly, int nlow, int nhigh,
                          double npeak, double sol_fraction, long long rseed, double tcorr,
                          double dt_update, int updates, double* out) {
  if (!modes || n_modes < 1 || updates < 0 || (updates > 0 && !out)) return JX_EINVAL;
  if (!(nlow <= npeak && npeak <= nhigh) || !(0.0 <= sol_fraction && sol_fraction <= 1.0))
    return JX_EINVAL;
  if (!std::isfinite(tcorr) || tcorr < 0.0 || !(dt_update > 0.0) || !std::isfinite(dt_update))
    return JX_EINVAL;
  const double dkx = 2.0 * M_PI / lx, dky = 2.0 * M_PI / ly;
  const double kpeak = npeak * 2.0 * M_PI / lx;
  const double band = nhigh * std::max(dkx, dky) - nlow * std::min(dkx, dky);
  const double pre = -4.0 / (band * band);
  std::vector<double> kx(n_modes), ky(n_modes), kiso(n_modes), norm(n_modes);
  for (int m = 0; m < n_modes; ++m) {
    kx[m] = modes[2 * m] * dkx;
    ky[m] = modes[2 * m + 1] * dky;
    kiso[m] = std::sqrt(kx[m] * kx[m] + ky[m] * ky[m]);
    const double off = kiso[m] - kpeak;
    const double power = std::fma(pre, off * off, 1.0);
    norm[m] = std::sqrt(std::fabs(power));
    norm[m] *= std::sqrt(kpeak / kiso[m]);
  }
  const double corr = tcorr > 1e-6 ? std::exp(-dt_update / tcorr) : 0.0;
  const double nw = std::sqrt(1.0 - corr * corr);
  Gaussian g(rseed);
  std::vector<double> noise(4 * (size_t)n_modes);
  for (int u = 0; u < updates; ++u) {
    for (auto& v : noise) v = g.next();
    for (int m = 0; m < n_modes; ++m) {
      double* q = &noise[4 * (size_t)m];   // [dir][re/im]
      for (int k = 0; k < 4; ++k) q[k] *= norm[m];
      const double k2 = kiso[m] * kiso[m];
      double res[4];
      for (int ri = 0; ri < 2; ++ri) {
        const double kdn = kx[m] * q[ri] + ky[m] * q[2 + ri];
        const double lon_x = kx[m] * kdn / k2, lon_y = ky[m] * kdn / k2;
        res[ri] = sol_fraction * (q[ri] - lon_x) + (1.0 - sol_fraction) * lon_x;
        res[2 + ri] = sol_fraction * (q[2 + ri] - lon_y) + (1.0 - sol_fraction) * lon_y;
      }
      double* dst = out + ((size_t)u * n_modes + m) * 4;
      const double* prev = (u == 0) ? nullptr : out + ((size_t)(u - 1) * n_modes + m) * 4;
      for (int k = 0; k < 4; ++k) dst[k] = (u == 0) ? res[k] : corr * prev[k] + nw * res[k];
    }
  }
  return JX_OK;
}

}  // extern "C"

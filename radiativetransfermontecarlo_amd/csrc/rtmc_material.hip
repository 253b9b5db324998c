// rtmc_material.hip -- rtmc_prepare_material: the scattering operator of
// MaterialClass (cross-sections, total cross-sections, inverse-CDF tables of
// the scattering angle) computed natively on the host and uploaded for the
// kernels.  Restates MaterialClass.m:172-359 (calcSigmaIsotropic),
// :381-710 (PSDF laws), :2379-2463 (prepareSigma, prepareSigmaOne), and the
// solid/air energy coefficients of elastic reflective boundaries
// (MaterialClass.m:1786-1830 with ZoeppritzFluid.m).
#include "rtmc_common.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace rtmc {
namespace {

constexpr double kPiH = 3.141592653589793;

// H(0) = 1/2 (MaterialClass.heaviside, MaterialClass.m:1760-1764)
double heaviside(double h) { return h > 0.0 ? 1.0 : (h == 0.0 ? 0.5 : 0.0); }

// Power spectral density Phi(z) of the correlation law, z = wavenumber * lc
// (MaterialClass.m:479-710).
struct Psdf {
  int law, d;
  double a, nu;
  double operator()(double z) const {
    switch (law) {
      case RTMC_PSDF_EXP:
        if (d == 2) return 1.0 / (4.0 * kPiH * std::pow(1.0 + (z / a) * (z / a), 1.5));
        return 1.0 / (6.0 * (kPiH * kPiH) * std::pow(1.0 + (z / a) * (z / a), 2.0));
      case RTMC_PSDF_POWER_LAW:
        if (d == 2) return z * std::cyl_bessel_k(1.0, z) / (4.0 * kPiH);
        return 1.0 / (6.0 * (kPiH * kPiH)) * std::exp(-z / a);
      case RTMC_PSDF_GAUSSIAN:
        if (d == 2) return 1.0 / (4.0 * kPiH) * std::exp(-(z * z) / 4.0);
        return 1.0 / (6.0 * (kPiH * kPiH)) * std::exp(-(z * z) / (4.0 * a));
      case RTMC_PSDF_TRIANGULAR:
        return (3.0 / 8.0 / std::pow(kPiH, 4.0)) * (1.0 - z / 2.0 / kPiH) * heaviside(2.0 * kPiH - z);
      case RTMC_PSDF_LOW_PASS:
        return (2.0 / 9.0 / std::pow(kPiH, 4.0)) * heaviside(a - z);
      default:  // von Karman
        if (d == 2) return 1.0 / (4.0 * kPiH) / std::pow(1.0 + (z / a) * (z / a), nu + 1.0);
        return 1.0 / (6.0 * (kPiH * kPiH)) / std::pow(1.0 + (z / a) * (z / a), nu + 1.5);
    }
  }
};

int make_psdf(const rtmc_material_params& p, Psdf* out) {
  Psdf P;
  P.law = p.spectral_law;
  P.d = p.dimension;
  P.nu = p.nu;
  P.a = 0.0;
  switch (p.spectral_law) {
    case RTMC_PSDF_EXP:
      P.a = p.dimension == 2 ? std::sqrt(2.0) : std::pow(6.0, 1.0 / 3.0);
      break;
    case RTMC_PSDF_POWER_LAW:
      P.a = std::pow(3.0 * kPiH / 4.0, 1.0 / 3.0);
      break;
    case RTMC_PSDF_GAUSSIAN:
      P.a = std::pow(3.0 * std::sqrt(kPiH) / 4.0, 2.0 / 3.0);
      break;
    case RTMC_PSDF_TRIANGULAR:
      RTMC_REQUIRE(p.dimension == 3, "dimensions other than 3 are not coded yet!");
      break;
    case RTMC_PSDF_LOW_PASS:
      RTMC_REQUIRE(p.dimension == 3, "LowPass spectral model is currently implemented only for d = 3.");
      P.a = 3.0 * kPiH / 2.0;
      break;
    case RTMC_PSDF_VON_KARMAN:
      RTMC_REQUIRE(p.nu > 0.0, "Please add the Hurst number for VonKarman PSDF");
      P.a = p.dimension == 2 ? 2.0 * std::sqrt(p.nu)
                             : std::pow(6.0 * std::sqrt(kPiH) * std::tgamma(p.nu + 1.5) / std::tgamma(p.nu), 1.0 / 3.0);
      break;
    default:
      set_error("unknown spectral law %d", p.spectral_law);
      return RTMC_ERR_ARG;
  }
  *out = P;
  return RTMC_OK;
}

// Differential scattering cross-sections (calcSigmaIsotropic, MaterialClass.m:172-359):
// channel 0 (acoustic) or 0..3 = PP, PS, SP, SS (elastic).
struct Sigma {
  Psdf phi;
  int d, acoustics, channel;
  double omega, zeta, zP, zS, K;
  double dkk, drr, rkr;            // acoustic
  double dll, dmm, rlm, rlr, rmr;  // elastic (drr shared)
  double operator()(double th) const {
    const double c = std::cos(th);
    if (acoustics) {
      const double p = d == 2 ? 2.0 : 3.0;
      return (kPiH / 2.0) * omega * std::pow(zeta, p) * (c * c * (drr * drr) + 2.0 * c * rkr * dkk * drr + dkk * dkk) *
             phi(zeta * std::sqrt(2.0 * (1.0 - c)));
    }
    const double K2 = K * K, K4 = std::pow(K, 4.0);
    if (d == 2) {
      switch (channel) {
        case 0:
          return (kPiH / 2.0) * omega * (zP * zP) *
                 ((1.0 - 2.0 / K2) * (1.0 - 2.0 / K2) * (dll * dll) + 4.0 * (1.0 / K2 - 2.0 / K4) * rlm * dll * dmm * (c * c) +
                  (4.0 / K4) * (dmm * dmm) * std::pow(c, 4.0) + (drr * drr) * (c * c) -
                  2.0 * (1.0 - 2.0 / K2) * rlr * dll * drr * c - (4.0 / K2) * rmr * dmm * drr * std::pow(c, 3.0)) *
                 phi(zP * std::sqrt(2.0 * (1.0 - c)));
        case 1:
          return (kPiH / 2.0) * omega * (zP * zP) *
                 (K2 * (drr * drr) + 4.0 * (dmm * dmm) * (c * c) - 4.0 * K * rmr * dmm * drr * c) * (1.0 - c * c) *
                 phi(zP * std::sqrt(1.0 + K2 - 2.0 * K * c));
        case 2:
          return (kPiH / 2.0) * omega * (zP * zP) *
                 ((drr * drr) + (4.0 / K2) * (dmm * dmm) * (c * c) - (4.0 / K) * rmr * dmm * drr * c) * (1.0 - c * c) *
                 phi(zS * std::sqrt(1.0 + 1.0 / K2 - 2.0 * c / K));
        default: {
          const double G = 2.0 * (c * c) - 1.0;
          return (kPiH / 2.0) * omega * (zS * zS) *
                 ((drr * drr) * (c * c) + (dmm * dmm) * (G * G) - 2.0 * rmr * dmm * drr * c * G) *
                 phi(zS * std::sqrt(2.0 * (1.0 - c)));
        }
      }
    }
    switch (channel) {
      case 0:
        return (kPiH / 2.0) * omega * std::pow(zP, 3.0) *
               ((1.0 - 2.0 / K2) * (1.0 - 2.0 / K2) * (dll * dll) + 4.0 * (1.0 / K2 - 2.0 / K4) * rlm * dll * dmm * (c * c) +
                (4.0 / K4) * (dmm * dmm) * std::pow(c, 4.0) + (drr * drr) * (c * c) -
                2.0 * (1.0 - 2.0 / K2) * rlr * dll * drr * c - (4.0 / K2) * rmr * dmm * drr * std::pow(c, 3.0)) *
               phi(zP * std::sqrt(2.0 * (1.0 - c)));
      case 1:
        return (kPiH / 2.0) * K * omega * std::pow(zP, 3.0) *
               (K2 * (drr * drr) + 4.0 * (dmm * dmm) * (c * c) - 4.0 * K * rmr * dmm * drr * c) * (1.0 - c * c) *
               phi(zP * std::sqrt(1.0 + K2 - 2.0 * K * c));
      case 2:
        return (kPiH / 4.0 / std::pow(K, 3.0)) * omega * std::pow(zS, 3.0) *
               ((drr * drr) + (4.0 / K2) * (dmm * dmm) * (c * c) - (4.0 / K) * rmr * dmm * drr * c) * (1.0 - c * c) *
               phi(zS * std::sqrt(1.0 + 1.0 / K2 - 2.0 / K * c));
      default: {
        const double ph = phi(zS * std::sqrt(2.0 * (1.0 - c)));
        const double tt = (kPiH / 4.0) * omega * std::pow(zS, 3.0) * (drr * drr) * (1.0 + c * c) * ph;
        const double gg = (kPiH / 4.0) * omega * std::pow(zS, 3.0) * (dmm * dmm) *
                          (4.0 * std::pow(c, 4.0) - 3.0 * (c * c) + 1.0) * ph;
        const double gt = -(kPiH / 4.0) * omega * std::pow(zS, 3.0) * rmr * dmm * drr * (4.0 * std::pow(c, 3.0)) * ph;
        return tt + gg + gt;
      }
    }
  }
};

// Gauss-Legendre, 20 nodes on [-1, 1] (positive half; the rule is symmetric)
const double kGLx[10] = {0.07652652113349734, 0.2277858511416451, 0.37370608871541955, 0.5108670019508271,
                         0.636053680726515,   0.7463319064601508, 0.8391169718222188,  0.9122344282513258,
                         0.9639719272779138,  0.9931285991850949};
const double kGLw[10] = {0.15275338713072578, 0.14917298647260366, 0.14209610931838187, 0.13168863844917653,
                         0.11819453196151825, 0.10193011981724026, 0.08327674157670467, 0.06267204833410944,
                         0.04060142980038622, 0.017614007139153273};

template <class F>
double gl20(const F& f, double a, double b) {
  const double m = 0.5 * (a + b), h = 0.5 * (b - a);
  double s = 0.0;
  for (int i = 0; i < 10; ++i) s += kGLw[i] * (f(m - h * kGLx[i]) + f(m + h * kGLx[i]));
  return s * h;
}

// Adaptive bisection of 20-point Gauss-Legendre: an interval is accepted when
// its rule and the sum over its halves agree to 1e-14 of the total (the
// integral of MaterialClass.m, at a tighter tolerance than MATLAB's 1e-6).
template <class F>
double integral(const F& f, double a, double b) {
  const int n0 = 16;
  double est = 0.0;
  struct Iv {
    double a, b, v;
    int depth;
  };
  std::vector<Iv> stack;
  for (int i = 0; i < n0; ++i) {
    double x0 = a + (b - a) * i / n0, x1 = a + (b - a) * (i + 1) / n0;
    double v = gl20(f, x0, x1);
    est += v;
    stack.push_back({x0, x1, v, 0});
  }
  const double tol = 1e-14 * std::fabs(est) + 1e-300;
  double total = 0.0;
  while (!stack.empty()) {
    Iv iv = stack.back();
    stack.pop_back();
    const double m = 0.5 * (iv.a + iv.b);
    const double l = gl20(f, iv.a, m), r = gl20(f, m, iv.b);
    if (std::fabs(l + r - iv.v) <= tol || iv.depth >= 50) {
      total += l + r;
    } else {
      stack.push_back({iv.a, m, l, iv.depth + 1});
      stack.push_back({m, iv.b, r, iv.depth + 1});
    }
  }
  return total;
}

// prepareSigmaOne (MaterialClass.m:2443-2463): Sigma, Sigma' and the inverse
// CDF of the scattering angle on n points of [0, pi].
void prepare_one(const Sigma& sg, int d, int64_t n, double thr, unsigned threads, double* S, double* Sp,
                 std::vector<double>* kx, std::vector<double>* ky) {
  if (d == 2) {
    *S = 2.0 * integral(sg, 0.0, kPiH);
    *Sp = 2.0 * integral([&](double t) { return sg(t) * std::cos(t); }, 0.0, kPiH);
  } else {
    *S = 2.0 * kPiH * integral([&](double t) { return sg(t) * std::sin(t); }, 0.0, kPiH);
    *Sp = 2.0 * kPiH * integral([&](double t) { return sg(t) * std::sin(t) * std::cos(t); }, 0.0, kPiH);
  }
  const double step = kPiH / (double)(n - 1);
  std::vector<double> cdf((size_t)n);
  // pdf on the grid in parallel chunks, then the running sum in order (nancumsum)
  {
    const double Sv = *S;
    const unsigned nt = std::max(1u, threads);
    const int64_t per = (n + nt - 1) / nt;
    auto part = [&](unsigned t) {
      const int64_t i0 = (int64_t)t * per, i1 = std::min(n, i0 + per);
      for (int64_t i = i0; i < i1; ++i) {
        const double x = (i == n - 1) ? kPiH : (double)i * step;
        cdf[(size_t)i] = (d == 2) ? (2.0 / Sv) * sg(x) : (2.0 * kPiH / Sv) * std::sin(x) * sg(x);
      }
    };
    if (nt == 1) {
      part(0);
    } else {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nt; ++t) th.emplace_back(part, t);
      for (auto& t : th) t.join();
    }
  }
  double run = 0.0;
  const double dx = kPiH / (double)n;
  for (int64_t i = 0; i < n; ++i) {
    const double pdf = cdf[(size_t)i];
    if (!std::isnan(pdf)) run += pdf;
    cdf[(size_t)i] = run * dx;
  }
  // knots i where the cdf step before or after exceeds the threshold
  kx->clear();
  ky->clear();
  for (int64_t i = 0; i < n; ++i) {
    const bool after = i + 1 < n && (cdf[(size_t)i + 1] - cdf[(size_t)i]) > thr;
    const bool before = i > 0 && (cdf[(size_t)i] - cdf[(size_t)i - 1]) > thr;
    if (!(after || before)) continue;
    const double th = (i == n - 1) ? kPiH : (double)i * step;
    if (thr <= 0.0 && !kx->empty() && !(cdf[(size_t)i] > kx->back())) continue;  // repeated knot
    kx->push_back(cdf[(size_t)i]);
    ky->push_back(th);
  }
}

struct DeviceTables {
  std::vector<void*> ptrs;
};

// ---- solid/air interface (MaterialClass.Zoeppritz, MaterialClass.m:1786-1830) ----
// The fluid is air (vp 343, vs 0, rho 1, :1799-1801) and the incidence angles
// linspace(0, 90, n) degrees (:1804).  For each angle ZoeppritzFluid.m solves the
// boundary conditions for P incidence (:49-121) and SV incidence (:123-188): the
// three equations left when the fluid carries no S wave, in the unknowns
// (reflected P, reflected SV, transmitted P) amplitudes.  Angles past a critical
// one are complex (asin of a slowness above 1), so the system is complex.
constexpr double kAirVp = 343.0, kAirVs = 0.0, kAirRho = 1.0;
constexpr int kZoeAngles = 181;
using cplx = std::complex<double>;

cplx casin(double v) { return std::asin(cplx(v, 0.0)); }

// Solves the 3x3 system M a = f by elimination with partial pivoting; false if singular.
bool solve3(cplx M[3][3], cplx f[3], cplx a[3]) {
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    for (int r = c + 1; r < 3; ++r)
      if (std::abs(M[r][c]) > std::abs(M[piv][c])) piv = r;
    if (std::abs(M[piv][c]) == 0.0) return false;
    if (piv != c) {
      for (int k = 0; k < 3; ++k) std::swap(M[c][k], M[piv][k]);
      std::swap(f[c], f[piv]);
    }
    for (int r = c + 1; r < 3; ++r) {
      const cplx m = M[r][c] / M[c][c];
      for (int k = c; k < 3; ++k) M[r][k] -= m * M[c][k];
      f[r] -= m * f[c];
    }
  }
  for (int r = 2; r >= 0; --r) {
    cplx acc = f[r];
    for (int k = r + 1; k < 3; ++k) acc -= M[r][k] * a[k];
    a[r] = acc / M[r][r];
  }
  return true;
}

// Rows of ZoeppritzFluid.m's boundary conditions that involve no fluid S wave,
// for angles phi1 (reflected P), psi1 (reflected S), phi2 (transmitted P),
// psi2 (transmitted S, zero slowness in the fluid).
void interface_matrix(double vp1, double vs1, double rho1, cplx phi1, cplx psi1, cplx phi2, cplx psi2,
                      cplx M[3][3]) {
  const double vp2 = kAirVp, vs2 = kAirVs, rho2 = kAirRho;
  M[0][0] = std::cos(phi1);
  M[0][1] = -std::sin(psi1);
  M[0][2] = std::cos(phi2);
  M[1][0] = -std::cos(2.0 * psi1);
  M[1][1] = vs1 / vp1 * std::sin(2.0 * psi1);
  M[1][2] = (rho2 * vp2 / (rho1 * vp1)) * std::cos(2.0 * psi2);
  M[2][0] = std::sin(2.0 * phi1);
  M[2][1] = vp1 / vs1 * std::cos(2.0 * psi1);
  M[2][2] = (rho2 * vp1 * vs2 * vs2 / (rho1 * vp2 * vs1 * vs1)) * std::sin(2.0 * phi2);
}

// E_Rpp (P incidence, reflected P energy) and E_Rsvsv (SV incidence, reflected
// SV energy) at n angles linspace(0, 90, n); NaN where a system is singular.
void zoeppritz_air(double vp1, double vs1, double rho1, int64_t n, double* deg, double* Rpp, double* Rsvsv) {
  const double vp2 = kAirVp, vs2 = kAirVs;
  for (int64_t i = 0; i < n; ++i) {
    // linspace(0, 90, n) as MATLAB forms it: d1 + (0:n-1) * (d2 - d1) / (n - 1), last = d2
    const double th = (i == n - 1) ? 90.0 : (double)i * (90.0 / (double)(n - 1));
    deg[i] = th;
    const double t = th * kPiH / 180.0;
    cplx M[3][3], f[3], a[3];
    {  // P incidence
      const double p = std::sin(t) / vp1;
      interface_matrix(vp1, vs1, rho1, cplx(t, 0.0), casin(vs1 * p), casin(vp2 * p), casin(vs2 * p), M);
      const cplx psi0 = casin(vs1 * std::sin(t) / vp1);
      f[0] = std::cos(t);
      f[1] = std::cos(2.0 * psi0);
      f[2] = std::sin(2.0 * t);
      Rpp[i] = solve3(M, f, a) ? std::norm(a[0]) : NAN;
    }
    {  // SV incidence
      const double p = std::sin(t) / vs1;
      interface_matrix(vp1, vs1, rho1, casin(vp1 * p), cplx(t, 0.0), casin(vp2 * p), casin(vs2 * p), M);
      f[0] = -std::sin(t);
      f[1] = -vs1 * std::sin(2.0 * t) / vp1;
      f[2] = vp1 * std::cos(2.0 * t) / vs1;
      Rsvsv[i] = solve3(M, f, a) ? std::norm(a[1]) : NAN;
    }
  }
}

}  // namespace
}  // namespace rtmc

using namespace rtmc;

namespace rtmc {
namespace {

unsigned host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// The material's cross-section functor for channel 0 (acoustic) with the
// frequency, velocities and fluctuation statistics of `p` (calcSigmaIsotropic).
int make_sigma(const rtmc_material_params* p, Sigma* out) {
  Psdf phi;
  int e = make_psdf(*p, &phi);
  if (e) return e;
  Sigma base;
  base.phi = phi;
  base.d = p->dimension;
  base.acoustics = p->acoustics;
  base.channel = 0;
  base.omega = 2.0 * kPiH * p->frequency;
  if (p->acoustics) {
    base.zeta = base.omega / p->v[0] * p->lc;  // MaterialClass.m:204
    base.dkk = p->cv[0];
    base.drr = p->cv[1];
    base.rkr = p->corr[0];
  } else {
    base.K = p->v[0] / p->v[1];  // :242-253
    base.zP = base.omega / p->v[0] * p->lc;
    base.zS = base.K * base.zP;
    base.dll = p->cv[0];
    base.dmm = p->cv[1];
    base.drr = p->cv[2];
    base.rlm = p->corr[0];
    base.rlr = p->corr[1];
    base.rmr = p->corr[2];
  }
  *out = base;
  return RTMC_OK;
}

int check_params(const rtmc_material_params* p) {
  RTMC_REQUIRE(p->dimension == 2 || p->dimension == 3, "dimension must be 2 or 3 (sigma for 1D not implemented)");
  RTMC_REQUIRE(p->frequency > 0.0 && p->lc > 0.0, "frequency and correlation length must be > 0");
  RTMC_REQUIRE(p->acoustics ? p->v[0] > 0.0 : (p->v[0] > 0.0 && p->v[1] > 0.0), "velocities must be > 0");
  if (p->acoustics) {
    RTMC_REQUIRE(std::fabs(p->corr[0]) <= 1.0, "Absolute value of corr(kappa,rho) should be less than 1!");
  } else {
    for (int k = 0; k < 3; ++k)
      RTMC_REQUIRE(std::fabs(p->corr[k]) <= 1.0, "Absolute values of correlation coefficients should be less than 1!");
  }
  RTMC_REQUIRE(p->n_theta == 0 || p->n_theta >= 16, "n_theta must be >= 16");
  RTMC_REQUIRE(p->rho >= 0.0 && std::isfinite(p->rho), "rho must be >= 0 and finite (0: no Zoeppritz tables)");
  return RTMC_OK;
}

// rtmc_prepare_material_host with `threads` host threads (1: all on the caller's)
int prepare_host(const rtmc_material_params* p, rtmc_material_host* out, unsigned threads) {
  RTMC_REQUIRE(p != nullptr && out != nullptr, "params and out must be given");
  memset(out, 0, sizeof(*out));
  int e = check_params(p);
  if (e) return e;
  const int64_t n = p->n_theta > 0 ? p->n_theta : 1000000;
  Sigma base;
  e = make_sigma(p, &base);
  if (e) return e;
  const int nch = p->acoustics ? 1 : 4;
  std::vector<double> kx[4], ky[4];
  double S[4] = {0, 0, 0, 0}, Sp[4] = {0, 0, 0, 0};
  auto channel = [&](int c, unsigned t) {
    Sigma sg = base;
    sg.channel = c;
    prepare_one(sg, p->dimension, n, p->knot_threshold, t, &S[c], &Sp[c], &kx[c], &ky[c]);
  };
  if (threads <= 1 || nch == 1) {
    for (int c = 0; c < nch; ++c) channel(c, threads);
  } else {
    std::vector<std::thread> th;
    const unsigned per = std::max(1u, threads / (unsigned)nch);
    for (int c = 0; c < nch; ++c) th.emplace_back(channel, c, per);
    for (auto& t : th) t.join();
  }
  for (int c = 0; c < nch; ++c) {
    if (kx[c].size() < 2) {
      set_error("inverse CDF of channel %d has %zu knots", c, kx[c].size());
      return RTMC_ERR_TABLE;
    }
    out->Sigma[c] = S[c];
    out->Sigma_prime[c] = Sp[c];
    out->n_knots[c] = (int64_t)kx[c].size();
    out->x[c] = (double*)malloc(sizeof(double) * kx[c].size());
    out->y[c] = (double*)malloc(sizeof(double) * kx[c].size());
    if (!out->x[c] || !out->y[c]) {
      rtmc_free_material_host(out);
      set_error("out of host memory");
      return RTMC_ERR_NOMEM;
    }
    memcpy(out->x[c], kx[c].data(), sizeof(double) * kx[c].size());
    memcpy(out->y[c], ky[c].data(), sizeof(double) * kx[c].size());
  }
  out->extrap = p->knot_threshold <= 0.0 ? RTMC_EXTRAP_NEAREST : RTMC_EXTRAP_LINEAR;
  if (p->acoustics) {
    out->mft[0] = out->mft[1] = 1.0 / S[0];  // MaterialClass.m:2385
    out->P2P = out->S2S = 1.0;
  } else {
    out->mft[0] = 1.0 / (S[0] + S[1]);  // :2412
    out->mft[1] = 1.0 / (S[2] + S[3]);
    out->P2P = S[0] / (S[0] + S[1]);    // :2436-2437
    out->S2S = S[3] / (S[3] + S[2]);
  }
  return RTMC_OK;
}

// Runs f(i) for i in [0, n) on up to host_threads() workers; returns the
// status of the first failure (its message re-raised on the caller's thread).
template <class F>
int parallel_items(int n, F f) {
  const unsigned nw = std::max(1u, std::min(host_threads(), (unsigned)std::max(n, 1)));
  std::atomic<int> next(0);
  std::atomic<int> bad(-1);
  std::vector<int> rc((size_t)std::max(n, 1), RTMC_OK);
  std::vector<std::string> msg((size_t)std::max(n, 1));
  auto work = [&]() {
    for (int i = next++; i < n; i = next++) {
      if (bad.load() >= 0) break;
      clear_error();
      rc[(size_t)i] = f(i);
      if (rc[(size_t)i] != RTMC_OK) {
        msg[(size_t)i] = last_error();
        int expect = -1;
        bad.compare_exchange_strong(expect, i);
      }
    }
  };
  if (nw == 1) {
    work();
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nw; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
  }
  const int b = bad.load();
  if (b >= 0) {
    set_error("material %d: %s", b, msg[(size_t)b].c_str());
    return rc[(size_t)b];
  }
  return RTMC_OK;
}

// Knots of the table kept at tolerance tol: 0, 1, n-2, n-1 and, in between,
// the end of each longest run whose chord from the run's first knot passes
// within tol of every knot it skips.  The admissible chord slopes from knot s
// form an interval that each skipped knot narrows (a cone); the run ends
// where the next knot's own slope leaves it.
void select_knots(const double* x, const double* y, int64_t n, double tol, std::vector<int64_t>* idx) {
  idx->clear();
  if (n <= 4) {
    for (int64_t i = 0; i < n; ++i) idx->push_back(i);
    return;
  }
  idx->push_back(0);
  idx->push_back(1);
  int64_t s = 1, k = 2;
  double lo = -INFINITY, hi = INFINITY;
  while (k <= n - 2) {
    const double dx = x[k] - x[s];
    const double sl = (y[k] - y[s]) / dx;
    if (sl >= lo && sl <= hi) {
      lo = std::max(lo, (y[k] - tol - y[s]) / dx);
      hi = std::min(hi, (y[k] + tol - y[s]) / dx);
      ++k;
    } else {
      s = k - 1;  // > the previous s: knot s + 1 always fits
      idx->push_back(s);
      lo = -INFINITY;
      hi = INFINITY;
      k = s + 1;
    }
  }
  if (idx->back() != n - 2) idx->push_back(n - 2);
  idx->push_back(n - 1);
}

// rtmc_compress_table on host arrays (checked by the caller)
void compress_table(const double* x, const double* y, int64_t n, double tol, int max_knots,
                    std::vector<rtmc_lds_knot>* out, double* tol_used) {
  std::vector<int64_t> idx;
  select_knots(x, y, n, tol, &idx);
  while ((int64_t)idx.size() > max_knots) {
    tol = tol > 0.0 ? 2.0 * tol : 1e-9;
    select_knots(x, y, n, tol, &idx);
  }
  const size_t m = idx.size();
  out->resize(m);
  for (size_t j = 0; j < m; ++j) {
    const int64_t i = idx[j];
    rtmc_lds_knot& r = (*out)[j];
    r.u = x[i];
    r.theta = (float)y[i];
    r.slope = 0.0f;
    if (j + 1 < m) {
      const int64_t i2 = idx[j + 1];
      r.slope = (float)((y[i2] - y[i]) / (x[i2] - x[i]));
    }
  }
  *tol_used = tol;
}

int check_compress_args(const double* x, const double* y, int64_t n, double tol, int32_t max_knots) {
  RTMC_REQUIRE(x != nullptr && y != nullptr, "table arrays must be given");
  RTMC_REQUIRE(n >= 2, "the table needs >= 2 knots");
  RTMC_REQUIRE(tol >= 0.0 && std::isfinite(tol), "tol must be >= 0 and finite");
  RTMC_REQUIRE(max_knots >= 4 && max_knots <= 65535, "max_knots must be in [4, 65535]");
  for (int64_t i = 0; i < n; ++i) {
    RTMC_REQUIRE(std::isfinite(x[i]) && std::isfinite(y[i]), "table knot %lld is not finite", (long long)i);
    RTMC_REQUIRE(i == 0 || x[i] > x[i - 1], "table knots must be strictly increasing (knot %lld)", (long long)i);
  }
  return RTMC_OK;
}

// The compressed tables of one material from its host tables, appended to `blob`
// (knot records); lds[c].knots holds the record offset until the upload.
int compress_material(const double* const* xs, const double* const* ys, const int64_t* ns, int nch, int extrap,
                      double tol, int max_knots, rtmc_lds_table* lds, std::vector<rtmc_lds_knot>* blob) {
  for (int c = 0; c < nch; ++c) {
    int e = check_compress_args(xs[c], ys[c], ns[c], tol, max_knots);
    if (e) return e;
    std::vector<rtmc_lds_knot> k;
    double used = tol;
    compress_table(xs[c], ys[c], ns[c], tol, max_knots, &k, &used);
    lds[c].m = (int32_t)k.size();
    lds[c].extrap = extrap;
    lds[c].tol = used;
    lds[c].knots = (const rtmc_lds_knot*)(uintptr_t)blob->size();
    blob->insert(blob->end(), k.begin(), k.end());
  }
  return RTMC_OK;
}

}  // namespace
}  // namespace rtmc

extern "C" int rtmc_prepare_material_host(const rtmc_material_params* p, rtmc_material_host* out) {
  clear_error();
  return prepare_host(p, out, host_threads());
}

extern "C" void rtmc_free_material_host(rtmc_material_host* out) {
  if (!out) return;
  for (int c = 0; c < 4; ++c) {
    free(out->x[c]);
    free(out->y[c]);
    out->x[c] = out->y[c] = nullptr;
    out->n_knots[c] = 0;
  }
}

extern "C" int rtmc_prepare_material(const rtmc_material_params* p, rtmc_material* out, void** handle) {
  clear_error();
  RTMC_REQUIRE(out != nullptr && handle != nullptr, "out and handle must be given");
  *handle = nullptr;
  rtmc_material_host h;
  int e = rtmc_prepare_material_host(p, &h);
  if (e) return e;
  DeviceTables* dt = new DeviceTables();
  memset(out, 0, sizeof(*out));
  out->acoustics = p->acoustics;
  out->vp = p->v[0];
  out->vs = p->acoustics ? 0.0 : p->v[1];
  out->mft[0] = h.mft[0];
  out->mft[1] = h.mft[1];
  out->P2P = h.P2P;
  out->S2S = h.S2S;
  out->QP = p->Q[0];
  out->QS = p->acoustics ? p->Q[0] : p->Q[1];
  out->frequency = p->frequency;
  const int nch = p->acoustics ? 1 : 4;
  int rc = RTMC_OK;
  for (int c = 0; c < nch && rc == RTMC_OK; ++c) {
    const size_t bytes = sizeof(double) * (size_t)h.n_knots[c];
    void* dx = nullptr;
    void* dy = nullptr;
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
        hipMemcpy(dx, h.x[c], bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy, h.y[c], bytes, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("uploading the inverse-CDF table failed");
      rc = RTMC_ERR_NOMEM;
    }
    if (dx) dt->ptrs.push_back(dx);
    if (dy) dt->ptrs.push_back(dy);
    out->invcdf[c].n = h.n_knots[c];
    out->invcdf[c].x = (const double*)dx;
    out->invcdf[c].y = (const double*)dy;
    out->invcdf[c].extrap = h.extrap;
  }
  if (rc == RTMC_OK && p->lds_tol > 0.0) {
    // compressed copies of the tables for the tracing kernel's LDS
    std::vector<rtmc_lds_knot> blob;
    const double* xs[4] = {h.x[0], h.x[1], h.x[2], h.x[3]};
    const double* ys[4] = {h.y[0], h.y[1], h.y[2], h.y[3]};
    rc = compress_material(xs, ys, h.n_knots, nch, h.extrap, p->lds_tol,
                           p->lds_max_knots > 0 ? p->lds_max_knots : 1024, out->lds, &blob);
    void* dk = nullptr;
    if (rc == RTMC_OK && (hipMalloc(&dk, blob.size() * sizeof(rtmc_lds_knot)) != hipSuccess ||
                          hipMemcpy(dk, blob.data(), blob.size() * sizeof(rtmc_lds_knot), hipMemcpyHostToDevice) !=
                              hipSuccess)) {
      set_error("uploading the compressed inverse-CDF tables failed");
      rc = RTMC_ERR_NOMEM;
    }
    if (dk) dt->ptrs.push_back(dk);
    for (int c = 0; c < nch && rc == RTMC_OK; ++c)
      out->lds[c].knots = (const rtmc_lds_knot*)dk + (uintptr_t)out->lds[c].knots;
  }
  rtmc_free_material_host(&h);
  if (rc == RTMC_OK && !p->acoustics && p->rho > 0.0) {
    // solid/air energy coefficients for elastic reflective boundaries
    double zh[3 * kZoeAngles];
    zoeppritz_air(p->v[0], p->v[1], p->rho, kZoeAngles, zh, zh + kZoeAngles, zh + 2 * kZoeAngles);
    void* dz = nullptr;
    if (hipMalloc(&dz, sizeof(zh)) != hipSuccess || hipMemcpy(dz, zh, sizeof(zh), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("uploading the Zoeppritz tables failed");
      rc = RTMC_ERR_NOMEM;
    }
    if (dz) dt->ptrs.push_back(dz);
    out->n_zoe = kZoeAngles;
    out->zoe_deg = (const double*)dz;
    out->zoe_Rpp = (const double*)dz + kZoeAngles;
    out->zoe_Rsvsv = (const double*)dz + 2 * kZoeAngles;
  }
  if (rc != RTMC_OK) {
    for (void* q : dt->ptrs) (void)hipFree(q);
    delete dt;
    memset(out, 0, sizeof(*out));
    return rc;
  }
  *handle = dt;
  return RTMC_OK;
}

extern "C" int rtmc_prepare_materials(const rtmc_material_params* params, int32_t n, rtmc_material* out,
                                      void** handle) {
  clear_error();
  RTMC_REQUIRE(handle != nullptr, "handle must be given");
  *handle = nullptr;
  RTMC_REQUIRE(n >= 0 && (n == 0 || (params != nullptr && out != nullptr)), "params and out must be given");
  if (n == 0) return RTMC_OK;
  // host tables: one material per worker, the angle grid of each on its worker
  std::vector<rtmc_material_host> h((size_t)n);
  for (auto& x : h) memset(&x, 0, sizeof(x));
  auto free_all = [&]() {
    for (auto& x : h) rtmc_free_material_host(&x);
  };
  // compressed LDS copies of the tables (lds_tol > 0), per material
  std::vector<std::vector<rtmc_lds_knot>> lblob((size_t)n);
  std::vector<rtmc_lds_table> ltab((size_t)n * 4);
  int e = parallel_items(n, [&](int i) {
    const rtmc_material_params* p = &params[i];
    int r = prepare_host(p, &h[(size_t)i], 1);
    if (r || !(p->lds_tol > 0.0)) return r;
    const rtmc_material_host& hi = h[(size_t)i];
    const double* xs[4] = {hi.x[0], hi.x[1], hi.x[2], hi.x[3]};
    const double* ys[4] = {hi.y[0], hi.y[1], hi.y[2], hi.y[3]};
    return compress_material(xs, ys, hi.n_knots, p->acoustics ? 1 : 4, hi.extrap, p->lds_tol,
                             p->lds_max_knots > 0 ? p->lds_max_knots : 1024, &ltab[(size_t)i * 4],
                             &lblob[(size_t)i]);
  });
  if (e) {
    free_all();
    return e;
  }
  // one device block for every table: x then y of each channel, 8-byte aligned,
  // then the compressed knots (16-byte records, 16-byte aligned)
  size_t total = 0;
  for (int i = 0; i < n; ++i) {
    const int nch = params[i].acoustics ? 1 : 4;
    for (int c = 0; c < nch; ++c) total += 2 * (size_t)h[(size_t)i].n_knots[c];
    if (!params[i].acoustics && params[i].rho > 0.0) total += 3 * (size_t)kZoeAngles;
  }
  total += total & 1;
  const size_t lds_off = total;  // in doubles
  for (int i = 0; i < n; ++i) total += 2 * lblob[(size_t)i].size();
  void* dev = nullptr;
  if (hipMalloc(&dev, total * sizeof(double)) != hipSuccess) {
    free_all();
    set_error("hipMalloc of %zu bytes of material tables failed", total * sizeof(double));
    return RTMC_ERR_NOMEM;
  }
  size_t off = 0;
  bool copied = true;
  for (int i = 0; i < n; ++i) {
    const rtmc_material_params* p = &params[i];
    const rtmc_material_host& hi = h[(size_t)i];
    rtmc_material& m = out[i];
    memset(&m, 0, sizeof(m));
    m.acoustics = p->acoustics;
    m.vp = p->v[0];
    m.vs = p->acoustics ? 0.0 : p->v[1];
    m.mft[0] = hi.mft[0];
    m.mft[1] = hi.mft[1];
    m.P2P = hi.P2P;
    m.S2S = hi.S2S;
    m.QP = p->Q[0];
    m.QS = p->acoustics ? p->Q[0] : p->Q[1];
    m.frequency = p->frequency;
    const int nch = p->acoustics ? 1 : 4;
    for (int c = 0; c < nch; ++c) {
      const size_t k = (size_t)hi.n_knots[c];
      copied = copied && hipMemcpy((double*)dev + off, hi.x[c], k * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
      copied = copied && hipMemcpy((double*)dev + off + k, hi.y[c], k * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
      m.invcdf[c].n = (int64_t)k;
      m.invcdf[c].x = (const double*)dev + off;
      m.invcdf[c].y = (const double*)dev + off + k;
      m.invcdf[c].extrap = hi.extrap;
      off += 2 * k;
    }
    if (!p->acoustics && p->rho > 0.0) {
      double zh[3 * kZoeAngles];
      zoeppritz_air(p->v[0], p->v[1], p->rho, kZoeAngles, zh, zh + kZoeAngles, zh + 2 * kZoeAngles);
      copied = copied && hipMemcpy((double*)dev + off, zh, sizeof(zh), hipMemcpyHostToDevice) == hipSuccess;
      m.n_zoe = kZoeAngles;
      m.zoe_deg = (const double*)dev + off;
      m.zoe_Rpp = (const double*)dev + off + kZoeAngles;
      m.zoe_Rsvsv = (const double*)dev + off + 2 * kZoeAngles;
      off += 3 * (size_t)kZoeAngles;
    }
  }
  for (int i = 0, loff = 0; i < n; ++i) {
    const std::vector<rtmc_lds_knot>& b = lblob[(size_t)i];
    if (b.empty()) continue;
    rtmc_lds_knot* base = (rtmc_lds_knot*)((double*)dev + lds_off) + loff;
    copied = copied && hipMemcpy(base, b.data(), b.size() * sizeof(rtmc_lds_knot), hipMemcpyHostToDevice) == hipSuccess;
    for (int c = 0; c < (params[i].acoustics ? 1 : 4); ++c) {
      rtmc_lds_table t = ltab[(size_t)i * 4 + c];
      t.knots = base + (uintptr_t)t.knots;
      out[i].lds[c] = t;
    }
    loff += (int)b.size();
  }
  free_all();
  if (!copied) {
    (void)hipFree(dev);
    memset(out, 0, sizeof(rtmc_material) * (size_t)n);  // nothing points at the freed block
    set_error("uploading the material tables failed");
    return RTMC_ERR_NOMEM;
  }
  DeviceTables* dt = new DeviceTables();
  dt->ptrs.push_back(dev);
  *handle = dt;
  return RTMC_OK;
}

extern "C" int rtmc_lc_derivatives(const rtmc_material_params* params, int32_t n, int64_t n_theta, double rel,
                                   double* dlogsigma_dlc, double* dlogSigma_dlc) {
  clear_error();
  RTMC_REQUIRE(n >= 0 && (n == 0 || (params != nullptr && dlogsigma_dlc != nullptr && dlogSigma_dlc != nullptr)),
               "params and outputs must be given");
  RTMC_REQUIRE(n_theta >= 2, "n_theta must be >= 2");
  RTMC_REQUIRE(rel > 0.0 && rel < 0.5, "rel must be in (0, 0.5)");
  return parallel_items(n, [&](int i) -> int {
    const rtmc_material_params& p = params[i];
    RTMC_REQUIRE(p.acoustics, "lc derivatives are implemented for acoustic media");
    int e = check_params(&p);
    if (e) return e;
    const double h = rel * p.lc;
    rtmc_material_params pp = p, pm = p;
    pp.lc = p.lc + h;
    pm.lc = p.lc - h;
    Sigma sp, sm;
    e = make_sigma(&pp, &sp);
    if (e) return e;
    e = make_sigma(&pm, &sm);
    if (e) return e;
    double* g = dlogsigma_dlc + (size_t)i * (size_t)n_theta;
    for (int64_t k = 0; k < n_theta; ++k) {
      const double th = (k == n_theta - 1) ? kPiH : (double)k * (kPiH / (double)(n_theta - 1));
      const double a = sp(th), b = sm(th);
      g[k] = (a > 0.0 && b > 0.0) ? (std::log(a) - std::log(b)) / (2.0 * h) : 0.0;
    }
    // total cross-sections as prepareSigmaOne integrates them (MaterialClass.m:2443-2455)
    double Sp_, Sm_;
    if (p.dimension == 2) {
      Sp_ = 2.0 * integral(sp, 0.0, kPiH);
      Sm_ = 2.0 * integral(sm, 0.0, kPiH);
    } else {
      Sp_ = 2.0 * kPiH * integral([&](double t) { return sp(t) * std::sin(t); }, 0.0, kPiH);
      Sm_ = 2.0 * kPiH * integral([&](double t) { return sm(t) * std::sin(t); }, 0.0, kPiH);
    }
    dlogSigma_dlc[i] = (std::log(Sp_) - std::log(Sm_)) / (2.0 * h);
    return RTMC_OK;
  });
}

extern "C" int rtmc_zoeppritz_air(double vp, double vs, double rho, int64_t n, double* deg, double* E_Rpp,
                                  double* E_Rsvsv) {
  clear_error();
  RTMC_REQUIRE(deg != nullptr && E_Rpp != nullptr && E_Rsvsv != nullptr, "outputs must be given");
  RTMC_REQUIRE(n >= 2, "n must be >= 2");
  RTMC_REQUIRE(vp > 0.0 && vs > 0.0 && rho > 0.0 && std::isfinite(vp) && std::isfinite(vs) && std::isfinite(rho),
               "vp, vs and rho must be positive and finite");
  zoeppritz_air(vp, vs, rho, n, deg, E_Rpp, E_Rsvsv);
  return RTMC_OK;
}

extern "C" int rtmc_compress_table(const double* x, const double* y, int64_t n, double tol, int32_t max_knots,
                                   rtmc_lds_knot* out, int32_t* m, double* tol_used) {
  clear_error();
  RTMC_REQUIRE(out != nullptr && m != nullptr && tol_used != nullptr, "outputs must be given");
  int e = check_compress_args(x, y, n, tol, max_knots);
  if (e) return e;
  std::vector<rtmc_lds_knot> k;
  compress_table(x, y, n, tol, max_knots, &k, tol_used);
  memcpy(out, k.data(), k.size() * sizeof(rtmc_lds_knot));
  *m = (int32_t)k.size();
  return RTMC_OK;
}

extern "C" int rtmc_attach_lds_tables(rtmc_material* mats, int32_t n, double tol, int32_t max_knots, void** handle,
                                      void* stream) {
  clear_error();
  RTMC_REQUIRE(handle != nullptr, "handle must be given");
  *handle = nullptr;
  RTMC_REQUIRE(n >= 0 && (n == 0 || mats != nullptr), "materials must be given");
  RTMC_REQUIRE(tol >= 0.0 && std::isfinite(tol), "tol must be >= 0 and finite");
  if (max_knots == 0) max_knots = 1024;
  RTMC_REQUIRE(max_knots >= 4 && max_knots <= 65535, "max_knots must be in [4, 65535]");
  if (n == 0) return RTMC_OK;
  for (int i = 0; i < n; ++i) {
    const int nch = mats[i].acoustics ? 1 : 4;
    for (int c = 0; c < nch; ++c) {
      const rtmc_table& t = mats[i].invcdf[c];
      RTMC_REQUIRE(t.n >= 2 && t.x != nullptr && t.y != nullptr, "material %d: inverse-CDF table %d is missing", i, c);
    }
  }
  // the tables may still be written by work queued on `stream`
  RTMC_HIP(hipStreamSynchronize((hipStream_t)stream));
  std::vector<std::vector<rtmc_lds_knot>> blob((size_t)n);
  std::vector<rtmc_lds_table> tab((size_t)n * 4);
  int e = parallel_items(n, [&](int i) -> int {
    const rtmc_material& M = mats[i];
    const int nch = M.acoustics ? 1 : 4;
    std::vector<double> hx[4], hy[4];
    const double* xs[4] = {nullptr, nullptr, nullptr, nullptr};
    const double* ys[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t ns[4] = {0, 0, 0, 0};
    for (int c = 0; c < nch; ++c) {
      const rtmc_table& t = M.invcdf[c];
      hx[c].resize((size_t)t.n);
      hy[c].resize((size_t)t.n);
      RTMC_HIP(hipMemcpy(hx[c].data(), t.x, sizeof(double) * (size_t)t.n, hipMemcpyDeviceToHost));
      RTMC_HIP(hipMemcpy(hy[c].data(), t.y, sizeof(double) * (size_t)t.n, hipMemcpyDeviceToHost));
      xs[c] = hx[c].data();
      ys[c] = hy[c].data();
      ns[c] = t.n;
      RTMC_REQUIRE(t.extrap == M.invcdf[0].extrap, "the tables of one material must share their extrapolation");
    }
    return compress_material(xs, ys, ns, nch, M.invcdf[0].extrap, tol, max_knots, &tab[(size_t)i * 4],
                             &blob[(size_t)i]);
  });
  if (e) return e;
  size_t total = 0;
  for (auto& b : blob) total += b.size();
  void* dev = nullptr;
  if (hipMalloc(&dev, total * sizeof(rtmc_lds_knot)) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipMalloc of %zu bytes of compressed tables failed", total * sizeof(rtmc_lds_knot));
    return RTMC_ERR_NOMEM;
  }
  // upload everything first: the materials are changed only once nothing can fail
  std::vector<size_t> base((size_t)n, 0);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const std::vector<rtmc_lds_knot>& b = blob[(size_t)i];
    base[(size_t)i] = off;
    if (hipMemcpy((rtmc_lds_knot*)dev + off, b.data(), b.size() * sizeof(rtmc_lds_knot), hipMemcpyHostToDevice) !=
        hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(dev);
      set_error("uploading the compressed tables failed");
      return RTMC_ERR_HIP;
    }
    off += b.size();
  }
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < (mats[i].acoustics ? 1 : 4); ++c) {
      rtmc_lds_table t = tab[(size_t)i * 4 + c];
      t.knots = (const rtmc_lds_knot*)dev + base[(size_t)i] + (uintptr_t)t.knots;
      mats[i].lds[c] = t;
    }
  }
  DeviceTables* dt = new DeviceTables();
  dt->ptrs.push_back(dev);
  *handle = dt;
  return RTMC_OK;
}

extern "C" int rtmc_free_material(void* handle) {
  clear_error();
  if (!handle) return RTMC_OK;
  DeviceTables* dt = (DeviceTables*)handle;
  for (void* q : dt->ptrs) RTMC_HIP(hipFree(q));
  delete dt;
  return RTMC_OK;
}

// rtmc_device.hpp -- device building blocks shared by the librtmc kernels
// (gfx950).  Physics follows cottereau/RadiativeTransferMonteCarlo; each
// routine cites the reference lines it implements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rtmc.h"

namespace rtmc {

constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kEps = 2.220446049250313e-16;  // MATLAB eps

// ------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Multipliers and Weyl constants
// are those of the paper.
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  // The key is opaque here, so the key schedule (k + r W) is recomputed by
  // scalar adds at each call instead of being hoisted out of the callers' loops
  // as 20 loop-invariant registers (which spill and then cost a v_readlane per
  // round).  The asm emits no instruction.
  asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53 random bits -> (0, 1]
__device__ __forceinline__ double u53(uint32_t w0, uint32_t w1) {
  uint64_t b = ((((uint64_t)w1) << 32) | w0) >> 11;
  return (double)(b + 1) * (1.0 / 9007199254740992.0);
}

// 24 random bits -> (0, 1] in FP32
__device__ __forceinline__ float u24(uint32_t w) {
  return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Variates of the function-level entry points (rtmc_draws in rtmc.h).
struct SlotDraws {
  int kind;
  int n_slots;
  const double* replay;
  int64_t n_steps;
  uint32_t k0, k1, stream;
  int64_t n;

  __device__ __forceinline__ double operator()(int64_t i, int slot, int64_t step) const {
    if (kind == RTMC_DRAWS_REPLAY) return replay[(step * n_slots + slot) * n + i];
    U4 c;
    c.x = (uint32_t)((uint64_t)i & 0xffffffffu);
    c.y = (uint32_t)(((uint64_t)i >> 32) & 0xffffu) | ((uint32_t)slot << 16);
    c.z = (uint32_t)step;
    c.w = stream;
    U4 r = philox10(c, k0, k1);
    return u53(r.x, r.y);
  }
};

// Global-memory pointers: loads and atomics through them compile to global, not
// flat, instructions (a flat access counts against the LDS wait counter too, so
// it holds up the LDS and cross-lane operations that wait on that counter).
#define RTMC_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ RTMC_GLOBAL T* gl(T* p) {
  return (RTMC_GLOBAL T*)p;
}

// ------------------------------------------------------------------
// Small vector helpers, generic in the real type (double: the reference's
// arithmetic; float: the FP32 free-flight variant).  Sums are left to right,
// as MATLAB's sum(.,2).
template <typename T>
__device__ __forceinline__ void cross3(const T a[3], const T b[3], T c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename T>
__device__ __forceinline__ T sumsq3(const T a[3]) { return (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]; }
template <typename T>
__device__ __forceinline__ T dot3(const T a[3], const T b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
template <typename T>
__device__ __forceinline__ T sgn(T v) { return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0)); }
__device__ __forceinline__ double acosd(double c) { return acos(c) * 180.0 / kPi; }
__device__ __forceinline__ double acosd(float c) { return (double)(acosf(c) * (180.0f / 3.14159265f)); }
__device__ __forceinline__ void sc(double a, double* s, double* c) { sincos(a, s, c); }
__device__ __forceinline__ void sc(float a, float* s, float* c) { sincosf(a, s, c); }
// sin and cos of 2 pi u (the reduction by whole turns is exact)
__device__ __forceinline__ void sc2pi(double u, double* s, double* c) { sincospi(2.0 * u, s, c); }
__device__ __forceinline__ void sc2pi(float u, float* s, float* c) { sincospif(2.0f * u, s, c); }

// ------------------------------------------------------------------
// griddedInterpolant(x, y)(q): linear inside, linear or nearest outside
// (MaterialClass.m:2462, initializeParticle.m:201, interp1(...,'extrap')
// MaterialClass.m:1674).  Interval = largest k <= n-2 with x[k] <= q.
template <typename PD>
__device__ __forceinline__ double table_eval(PD __restrict__ x, PD __restrict__ y, int64_t n, int extrap, double q) {
  if (q != q) return q;
  int64_t k;
  if (q < x[0]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[0];
    k = 0;
  } else if (q > x[n - 1]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[n - 1];
    k = n - 2;
  } else {
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (x[mid] <= q) lo = mid;
      else hi = mid;
    }
    k = lo;
  }
  double t = (q - x[k]) / (x[k + 1] - x[k]);
  return y[k] + t * (y[k + 1] - y[k]);
}

// Same lookup, restricted to the knot range [lo, hi] known to bracket q
// (guide-table search: lo = largest knot <= q is inside [lo, hi]).
template <typename PD>
__device__ __forceinline__ double table_eval_bracket(PD __restrict__ x, PD __restrict__ y, int64_t n, int extrap,
                                                     double q, int64_t lo, int64_t hi) {
  if (q != q) return q;
  int64_t k;
  if (q < x[0]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[0];
    k = 0;
  } else if (q > x[n - 1]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[n - 1];
    k = n - 2;
  } else {
    if (hi > n - 1) hi = n - 1;
    if (lo < 0) lo = 0;
    if (lo > n - 2) lo = n - 2;
    // widen if the bracket does not hold (defensive; never needed for a valid guide)
    if (x[lo] > q) lo = 0;
    if (hi <= lo || !(x[hi] > q)) hi = n - 1;
    while (hi - lo > 1) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (x[mid] <= q) lo = mid;
      else hi = mid;
    }
    k = lo;
  }
  double t = (q - x[k]) / (x[k + 1] - x[k]);
  return y[k] + t * (y[k + 1] - y[k]);
}

// table_eval with the interval bracketed exactly, x[lo] <= q < x[hi] (a guide
// built from this non-decreasing x): no widening checks, so the knots are read
// once, after the few search steps (the bracket usually holds one interval).
template <typename PD>
__device__ __forceinline__ double table_eval_exact_bracket(PD __restrict__ x, PD __restrict__ y, int64_t n, int extrap,
                                                           double q, int64_t lo, int64_t hi) {
  if (q != q) return q;
  int64_t k;
  if (q < x[0]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[0];
    k = 0;
  } else if (q > x[n - 1]) {
    if (extrap == RTMC_EXTRAP_NEAREST) return y[n - 1];
    k = n - 2;
  } else {
    lo = lo < 0 ? 0 : (lo > n - 2 ? n - 2 : lo);
    hi = hi > n - 1 ? n - 1 : (hi <= lo ? lo + 1 : hi);
    while (hi - lo > 1) {
      int64_t mid = lo + ((hi - lo) >> 1);
      if (x[mid] <= q) lo = mid;
      else hi = mid;
    }
    k = lo;
  }
  const double x0 = x[k], x1 = x[k + 1], y0 = y[k], y1 = y[k + 1];
  double t = (q - x0) / (x1 - x0);
  return y0 + t * (y1 - y0);
}

struct TableDev {
  const double* x;
  const double* y;
  int64_t n;
  int extrap;
};

// ------------------------------------------------------------------
// Material as the kernels read it.
struct MatDev {
  int acoustics;
  double vp, vs;
  double mftP, mftS;
  double P2P, S2S;
  TableDev inv[4];
  const double* zdeg;
  const double* zRpp;
  const double* zRss;
  int64_t nz;
  // compressed tables for the tracing kernel's LDS (rtmc_material.lds; lm = 0: none)
  const rtmc_lds_knot* lk[4];
  int lm[4];
  int lnear;  // their extrapolation is RTMC_EXTRAP_NEAREST (the lookup clamps to the end knots)
};

// ------------------------------------------------------------------
// Boundary interaction of one particle, given that it crossed boundary B
// during the current sub-step (propagateParticleSmallDt.m:167-581).
// uA, uB: the two boundary variates (P: reflect-or-convert; S: SV/SH split;
// SV: reflect-or-convert).
// The boundary code indexes vectors only with constants and, where branches
// write different components, works on local copies written back once at the
// end: a run-time index, or stores to different addresses merged at a join,
// would keep the particle's vectors in the private segment.
template <typename T>
__device__ __forceinline__ T comp3(const T a[3], int b) {
  const T a0 = a[0], a1 = a[1], a2 = a[2];
  return b == 0 ? a0 : (b == 1 ? a1 : a2);
}

template <typename T>
__device__ __forceinline__ bool plane_crossed(const rtmc_boundary& B, const T x[3], const T x1[3]) {
  int b = B.dir - 1;
  const T v = (T)B.val;
  return (comp3(x, b) - v) * (comp3(x1, b) - v) < T(0);  // :171-172
}

template <typename T>
__device__ __forceinline__ bool cyl_outside(const rtmc_boundary& B, const T x[3]) {
  T r = sqrt(x[0] * x[0] + x[1] * x[1]);  // :328-329
  return (r - (T)B.val) > T(0);            // :332-334
}

// MaterialClass.MaterialInterface branch at plane of normal e_b
// (MaterialClass.m:1587-1735): 0 P->P, 1 P->SV, 2 SH, 3 SV->SV, 4 SV->P.
template <int b, typename T>
__device__ __forceinline__ int interface_plane(const MatDev& M, const T d[3], const T q[3], int isP, double uA,
                                               double uB) {
  T nv[3] = {T(0), T(0), T(0)};
  nv[b] = T(1);
  if (isP) {
    T a = fabs(dot3(d, nv));  // :1597-1599
    if (a > T(1)) a = T(1);
    double inc = acosd(a);                                                 // :1600
    double R = table_eval(M.zdeg, M.zRpp, M.nz, RTMC_EXTRAP_LINEAR, inc);  // :1679
    return (uA <= R) ? 0 : 1;                                              // :1681-1682
  }
  T eSH[3];
  cross3(d, nv, eSH);  // :1613
  T nsh = sqrt(sumsq3(eSH));
  if (nsh < (T)kEps) {  // :1616-1625
    T t[3] = {T(1), T(0), T(0)};
    T tn = dot3(t, nv);
    for (int k = 0; k < 3; ++k) t[k] = t[k] - tn * nv[k];
    for (int k = 0; k < 3; ++k) eSH[k] = t[k];
    nsh = sqrt(sumsq3(t));
  }
  for (int k = 0; k < 3; ++k) eSH[k] = eSH[k] / nsh;
  T eSV[3];
  cross3(eSH, d, eSV);  // :1630-1631
  T nsv = sqrt(sumsq3(eSV));
  for (int k = 0; k < 3; ++k) eSV[k] = eSV[k] / nsv;
  T aSV = dot3(q, eSV), aSH = dot3(q, eSH);  // :1634-1643
  T ESV = aSV * aSV, ESH = aSH * aSH;
  T Etot = ESV + ESH + (T)kEps;
  T probSV = ESV / Etot;
  if (!(uA < (double)probSV)) return 2;  // :1646, SH always reflects :1720
  T a = fabs(dot3(d, nv));
  if (a > T(1)) a = T(1);
  double inc = acosd(a);
  double R = table_eval(M.zdeg, M.zRss, M.nz, RTMC_EXTRAP_LINEAR, inc);  // :1700
  return (uB <= R) ? 3 : 4;                                              // :1702-1703
}

// Reflection at the plane x_b = val (acoustic :184-187; elastic :190-321).
template <int b, typename T>
__device__ __forceinline__ void plane_reflect_axis(const MatDev& M, bool acoustic, double val, T xo[3], T dout[3],
                                                   T qo[3], int& isPo, double uA, double uB) {
  T x[3] = {xo[0], xo[1], xo[2]}, d[3] = {dout[0], dout[1], dout[2]}, q[3] = {qo[0], qo[1], qo[2]};
  int isP = isPo;
  const T v2 = (T)(2 * val);
  if (acoustic) {
    x[b] = v2 - x[b];
    d[b] = -d[b];
  } else {
    int br = interface_plane<b>(M, d, q, isP, uA, uB);
    x[b] = v2 - x[b];  // :198
    if (br == 0 || br == 2 || br == 3) {
      d[b] = -d[b];
      q[b] = -q[b];
    } else if (br == 1) {  // P -> SV :212-248
      isP = 0;
      T ratio = (T)(M.vs / M.vp);
      T dn = d[b];
      T t[3] = {d[0], d[1], d[2]};
      t[b] = T(0);
      for (int k = 0; k < 3; ++k) t[k] = ratio * t[k];
      T sin2 = sumsq3(t);
      T cv = sqrt(fmax(T(0), T(1) - sin2));
      t[b] = -sgn(dn) * cv;
      T nn = sqrt(sumsq3(t));
      for (int k = 0; k < 3; ++k) d[k] = t[k] / nn;
      T nv[3] = {T(0), T(0), T(0)};
      nv[b] = T(1);
      T ax[3];
      cross3(d, nv, ax);
      T na = sqrt(sumsq3(ax));
      for (int k = 0; k < 3; ++k) ax[k] = ax[k] / na;
      cross3(ax, d, q);
    } else {  // SV -> P :274-320
      isP = 1;
      T ratio = (T)(M.vp / M.vs);
      T dn = d[b];
      T t[3] = {d[0], d[1], d[2]};
      t[b] = T(0);
      for (int k = 0; k < 3; ++k) t[k] = ratio * t[k];
      T sin2 = sumsq3(t);
      if (sin2 <= T(1)) {
        T cv = sqrt(T(1) - sin2);
        t[b] = -sgn(dn) * cv;
        T nn = sqrt(sumsq3(t));
        for (int k = 0; k < 3; ++k) d[k] = t[k] / nn;
      } else {
        isP = 0;  // total reflection fallback :313-319
        d[b] = -d[b];
      }
    }
  }
  for (int k = 0; k < 3; ++k) {
    xo[k] = x[k];
    dout[k] = d[k];
    qo[k] = q[k];
  }
  isPo = isP;
}

// Applies the reflection of a particle that crossed plane boundary B.
template <typename T>
__device__ __forceinline__ void plane_reflect(const MatDev& M, bool acoustic, const rtmc_boundary& B, T x[3], T d[3],
                                              T q[3], int& isP, double uA, double uB) {
  if (B.dir == 1)
    plane_reflect_axis<0>(M, acoustic, B.val, x, d, q, isP, uA, uB);
  else if (B.dir == 2)
    plane_reflect_axis<1>(M, acoustic, B.val, x, d, q, isP, uA, uB);
  else
    plane_reflect_axis<2>(M, acoustic, B.val, x, d, q, isP, uA, uB);
}

template <typename T>
__device__ __forceinline__ void reflect_dq(T d[3], T q[3], const T nv[3]) {
  T dtn = dot3(d, nv);
  T ptn = dot3(q, nv);
  for (int k = 0; k < 3; ++k) {
    d[k] = d[k] - T(2) * dtn * nv[k];
    q[k] = q[k] - T(2) * ptn * nv[k];
  }
}

// Applies the reflection of a particle outside cylinder B
// (propagateParticleSmallDt.m:344-578).
template <typename T>
__device__ __forceinline__ void cyl_reflect(const MatDev& M, bool acoustic, const rtmc_boundary& B, T xo[3],
                                            T dout[3], T qo[3], int& isPo, double uA, double uB) {
  T x[3] = {xo[0], xo[1], xo[2]}, d[3] = {dout[0], dout[1], dout[2]}, q[3] = {qo[0], qo[1], qo[2]};
  int isP = isPo;
  T r = sqrt(x[0] * x[0] + x[1] * x[1]);
  T ov = r - (T)B.val;
  T nv[3] = {x[0] / r, x[1] / r, T(0)};
  for (int k = 0; k < 3; ++k) x[k] = x[k] - T(2) * ov * nv[k];  // :352
  if (acoustic) {                                               // :354-359
    T dp = dot3(d, nv);
    for (int k = 0; k < 3; ++k) d[k] = d[k] - T(2) * dp * nv[k];
  } else if (isP) {  // :373-455
    T ct = dot3(d, nv);
    double thd = acosd(fmin(T(1), fmax(T(-1), ct)));
    double R = table_eval(M.zdeg, M.zRpp, M.nz, RTMC_EXTRAP_NEAREST, thd);
    if (uA <= R) {
      reflect_dq(d, q, nv);
    } else {
      isP = 0;
      T eta = (T)(M.vs / M.vp);
      T ddn = dot3(d, nv);
      T dt_[3], dnt[3];
      for (int k = 0; k < 3; ++k) {
        T dnv = ddn * nv[k];
        dt_[k] = d[k] - dnv;
      }
      for (int k = 0; k < 3; ++k) dnt[k] = eta * dt_[k];
      T sin2 = sumsq3(dnt);
      T cv = sqrt(fmax(T(0), T(1) - sin2));
      T dnew[3];
      for (int k = 0; k < 3; ++k) dnew[k] = dnt[k] - cv * nv[k];
      T nn = sqrt(sumsq3(dnew));
      for (int k = 0; k < 3; ++k) dnew[k] = dnew[k] / nn;
      for (int k = 0; k < 3; ++k) d[k] = dnew[k];
      T ax[3];
      cross3(dnew, nv, ax);
      T al = sqrt(sumsq3(ax));
      if (al < T(1e-6)) {
        T e1[3] = {T(1), T(0), T(0)};
        cross3(dnew, e1, ax);
      }
      T an = sqrt(sumsq3(ax));
      for (int k = 0; k < 3; ++k) ax[k] = ax[k] / an;
      cross3(ax, dnew, q);
    }
  } else {  // :458-576
    T ai[3];
    cross3(d, nv, ai);
    T an = sqrt(sumsq3(ai));
    for (int k = 0; k < 3; ++k) ai[k] = ai[k] / an;
    T shc = dot3(q, ai);
    T prob_sh = shc * shc;
    if (uA < (double)prob_sh) {
      reflect_dq(d, q, nv);
    } else {
      T ct = dot3(d, nv);
      double thd = acosd(fmin(T(1), fmax(T(-1), ct)));
      double R = table_eval(M.zdeg, M.zRss, M.nz, RTMC_EXTRAP_NEAREST, thd);
      if (uB <= R) {
        reflect_dq(d, q, nv);
      } else {
        isP = 1;
        T eta = (T)(M.vp / M.vs);
        T ddn = dot3(d, nv);
        T dt_[3], dnt[3];
        for (int k = 0; k < 3; ++k) {
          T dnv = ddn * nv[k];
          dt_[k] = d[k] - dnv;
        }
        for (int k = 0; k < 3; ++k) dnt[k] = eta * dt_[k];
        T sin2 = sumsq3(dnt);
        if (sin2 <= T(1)) {
          T cv = sqrt(T(1) - sin2);
          T dnew[3];
          for (int k = 0; k < 3; ++k) dnew[k] = dnt[k] - cv * nv[k];
          T nn = sqrt(sumsq3(dnew));
          for (int k = 0; k < 3; ++k) d[k] = dnew[k] / nn;
        } else {
          isP = 0;
          reflect_dq(d, q, nv);
        }
      }
    }
  }
  for (int k = 0; k < 3; ++k) {
    xo[k] = x[k];
    dout[k] = d[k];
    qo[k] = q[k];
  }
  isPo = isP;
}

// ------------------------------------------------------------------
// Scattering of one particle (propagateParticleSmallDt.m:604-639,
// propagateParticle.m:33-72).  bit2d: the 2-D randi([0 1]) sign bit;
// uphi: 3-D azimuth uniform; uth: polar-angle uniform; uconv: the P2P (P)
// or S2S (S) uniform.  Returns theta (from the double-precision table).
// FAST (the fused kernel): sin and cos of the azimuth by sincospi of 2 uphi
// instead of sincos of 2 pi uphi (the same angle, without the reduction).
template <int DIM, typename T, typename TabF, bool FAST = false>
__device__ __forceinline__ double scatter_angles_t(const MatDev& M, bool acoustic, T d[3], T q[3], int& isP,
                                                   bool bit2d, double uphi, double uth, double uconv, TabF lookup) {
  if (DIM == 2) {
    T s = bit2d ? T(-1) : T(1);
    for (int k = 0; k < 3; ++k) q[k] = s * q[k];
  } else {
    T c[3];
    cross3(d, q, c);
    T sp, cp;
    if (FAST) {
      sc2pi((T)uphi, &sp, &cp);
    } else {
      T phi = (T)(kTwoPi * uphi);
      sc(phi, &sp, &cp);
    }
    for (int k = 0; k < 3; ++k) q[k] = cp * q[k] + sp * c[k];
  }
  if (acoustic) return lookup(0, uth);
  int from_p = isP;
  if (from_p) {
    if (uconv > M.P2P) isP = 0;
  } else {
    if (uconv > M.S2S) isP = 1;
  }
  int tab = from_p ? (isP ? 0 : 1) : (isP ? 2 : 3);
  return lookup(tab, uth);
}

template <int DIM>
__device__ __forceinline__ double scatter_angles(const MatDev& M, double d[3], double q[3], int& isP, bool bit2d,
                                                 double uphi, double uth, double uconv) {
  return scatter_angles_t<DIM>(M, M.acoustics != 0, d, q, isP, bit2d, uphi, uth, uconv, [&](int t, double u) {
    const TableDev& T = M.inv[t];
    return table_eval(T.x, T.y, T.n, T.extrap, u);
  });
}

// dir <- cos(th) dir + sin(th) perp, perp <- -sin(th) dir + cos(th) perp,
// both renormalised (propagateParticleSmallDt.m:635-639).
template <typename T>
__device__ __forceinline__ void rotate_renorm(T d[3], T q[3], T th) {
  T st, ct;
  sc(th, &st, &ct);
  T nd[3], np[3];
  for (int k = 0; k < 3; ++k) nd[k] = ct * d[k] + st * q[k];
  T nrm = sqrt(sumsq3(nd));
  for (int k = 0; k < 3; ++k) nd[k] = nd[k] / nrm;
  for (int k = 0; k < 3; ++k) np[k] = -st * d[k] + ct * q[k];
  T nrp = sqrt(sumsq3(np));
  for (int k = 0; k < 3; ++k) np[k] = np[k] / nrp;
  for (int k = 0; k < 3; ++k) {
    d[k] = nd[k];
    q[k] = np[k];
  }
}

// 1/sqrt(s) for a normal s > 0: the hardware estimate (v_rsq_f64, good to
// about 2^-23 relative) refined by two Newton steps y <- y (3 - s y^2) / 2,
// each of which squares the relative error, so the result is within a few
// ulp of the rounded 1/sqrt(s), without the division and the correctly
// rounded square root.
__device__ __forceinline__ double rsqrt_nr(double s) {
  double y = __builtin_amdgcn_rsq(s);
  const double hs = 0.5 * s;
  y = y * fma(-hs, y * y, 1.5);
  y = y * fma(-hs, y * y, 1.5);
  return y;
}
// FP32: v_rsq_f32 (1 ulp) refined once
__device__ __forceinline__ float rsqrt_nr(float s) {
  float y = __builtin_amdgcn_rsqf(s);
  return y * fmaf(-0.5f * s, y * y, 1.5f);
}

// rotate_renorm for the fused kernel: sin and cos of theta through sincospi
// of theta/pi (the reduction by whole turns is exact; theta times the rounded
// 1/(2 pi) moves the angle by about 1e-15 rad), the norms from rsqrt_nr (the
// vectors are unit to a few ulp, so the result differs from the division by
// the correctly rounded norm by a few ulp)
template <typename T>
__device__ __forceinline__ void rotate_renorm_fast(T d[3], T q[3], T th) {
  T st, ct;
  sc2pi(th * (T)(0.5 / kPi), &st, &ct);
  T nd[3], np[3];
  for (int k = 0; k < 3; ++k) nd[k] = ct * d[k] + st * q[k];
  for (int k = 0; k < 3; ++k) np[k] = -st * d[k] + ct * q[k];
  const T id = rsqrt_nr(sumsq3(nd)), ip = rsqrt_nr(sumsq3(np));
  for (int k = 0; k < 3; ++k) {
    d[k] = nd[k] * id;
    q[k] = np[k] * ip;
  }
}

template <typename T>
__device__ __forceinline__ void renorm(T d[3], T q[3]) {
  T nrm = sqrt(sumsq3(d));
  for (int k = 0; k < 3; ++k) d[k] = d[k] / nrm;
  T nrp = sqrt(sumsq3(q));
  for (int k = 0; k < 3; ++k) q[k] = q[k] / nrp;
}

// ------------------------------------------------------------------
// Observation variables (observeTime.m:15-45): 0 x|r, 1 y|azimuth,
// 2 z|sin(elevation), 3 direction.  `need` masks the ones to compute.
// Computed in the flight precision T, returned in double for the binning.
// Radii are hypot(hypot(x, y), z) as MATLAB's cart2pol / cart2sph compute
// them; with FAST they are sqrt(x^2 + y^2 [+ z^2]), which differs by at most
// an ulp or two and so moves a deposit only when the radius lies that close to
// a bin edge (positions here are far from overflow or underflow).
// Direction variable with the direction's cosine against the position vector
// (RTMC_DIRVAR_POSITION, README.md:209): cos = d . X / |X|, NaN at the origin.
template <typename T>
__device__ __forceinline__ T cos_to_position(const T d[3], const T X[3]) {
  const T dot = d[0] * X[0] + d[1] * X[1] + d[2] * X[2];
  return dot / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
}

// dz: the direction's z component, or (RTMC_DIRVAR_POSITION) its cosine
// against the position vector (cos_to_position); clamped to [-1, 1] here.
template <typename T, bool FAST = false>
__device__ __forceinline__ void obs_vars(int dim, int frame, const T X[3], T dz, unsigned need, double v[4]) {
  T r = T(0), th = T(0), ph = T(0);
  bool polar = frame != RTMC_FRAME_CARTESIAN;
  auto r2 = [&](T a, T b) { return FAST ? sqrt(a * a + b * b) : hypot(a, b); };
  if (dim == 2) {
    if (polar && (need & 1u)) r = r2(X[0], X[1]);
    if (polar && (need & 2u)) th = atan2(X[1], X[0]);
  } else if (frame == RTMC_FRAME_CYLINDRICAL) {
    if (need & 1u) r = r2(X[0], X[1]);
    if (need & 2u) th = atan2(X[1], X[0]);
  } else if (frame == RTMC_FRAME_SPHERICAL) {
    T hxy = T(0);
    if (FAST) {
      if (need & 4u) hxy = sqrt(X[0] * X[0] + X[1] * X[1]);
      if (need & 1u) r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    } else {
      if (need & 5u) hxy = hypot(X[0], X[1]);
      if (need & 1u) r = hypot(hxy, X[2]);
    }
    if (need & 2u) th = atan2(X[1], X[0]);
    if (need & 4u) ph = sin(atan2(X[2], hxy));
  }
  T cp = dz;
  if (cp > T(1)) cp = T(1);
  if (cp < T(-1)) cp = T(-1);
  T psi = T(0);
  if (need & 8u) psi = (dim == 3) ? cp : acos(cp);
  if (frame == RTMC_FRAME_CARTESIAN) {
    v[0] = X[0];
    v[1] = X[1];
    v[2] = X[2];
  } else if (frame == RTMC_FRAME_CYLINDRICAL) {
    v[0] = r;
    v[1] = th;
    v[2] = X[2];
  } else {
    v[0] = r;
    v[1] = th;
    v[2] = ph;
  }
  v[3] = psi;
}

// histcounts bin: e_k <= v < e_k+1, last bin closed; -1 outside or NaN.
__device__ __forceinline__ int64_t bin_search(const double* __restrict__ e, int ne, double v) {
  if (!(v >= e[0] && v <= e[ne - 1])) return -1;
  int64_t lo = 0, hi = ne - 1;
  while (hi - lo > 1) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (e[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Same bin with a guess from the mean spacing, corrected against the true
// edges (exact for any non-decreasing edges; O(1) steps when near-uniform).
__device__ __forceinline__ int bin_guess(const double* __restrict__ e, int ne, double e0, double inv_w, double v) {
  if (!(v >= e0 && v <= e[ne - 1])) return -1;
  int nb = ne - 1;
  double g = (v - e0) * inv_w;
  int k = (g >= 0.0 && g < (double)nb) ? (int)g : (g >= (double)nb ? nb - 1 : 0);
  while (k > 0 && v < e[k]) --k;
  while (k < nb - 1 && v >= e[k + 1]) ++k;
  return k;
}

// bin_guess with the last edge `ehi` known: the guessed bin's two edges are read
// together and usually confirm it; otherwise the search walks as bin_guess does.
__device__ __forceinline__ int bin_guess2(const double* __restrict__ e, int ne, double e0, double ehi, double inv_w,
                                          double v) {
  if (!(v >= e0 && v <= ehi)) return -1;
  const int nb = ne - 1;
  const double g = (v - e0) * inv_w;
  int k = (g >= 0.0 && g < (double)nb) ? (int)g : (g >= (double)nb ? nb - 1 : 0);
  const double lo = e[k], hi = e[k + 1];
  if ((v >= lo) & ((v < hi) | (k == nb - 1))) return k;  // both edges read unconditionally (one LDS access)
  while (k > 0 && v < e[k]) --k;
  while (k < nb - 1 && v >= e[k + 1]) ++k;
  return k;
}

}  // namespace rtmc

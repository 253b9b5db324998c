// rtmc_trace.hpp -- the persistent tracing kernel behind rtmc_forward,
// rtmc_forward_batched and rtmc_adjoint_batched.
//
// Replaces the packet x time loop of radiativeTransfer.m:40-158.  One lane
// owns one particle from birth to the last observation time; there is no
// particle state in HBM, only the histogram.
//
// Sub-stepped process (material.timeSteps = 0).  propagateParticleSmallDt
// advances every particle by sub-steps of length dt_s and, in each sub-step:
// removes it with probability pKill (Q attenuation), moves it, applies the
// boundaries, and scatters it with probability 1 - exp(-dt_s/mft).  These
// Bernoulli trials are independent, so the index of the next sub-step with a
// scattering (or removal) is geometric; the kernel samples those indices
// directly (G = ceil(E / lambda), E ~ Exp(1)) and jumps from event to event,
// executing a sub-step in full only where something happens: a scattering, a
// removal, or a predicted boundary crossing.  In between the particle flies
// straight exactly as the sub-step loop would move it.  The sampled process
// has the same law as the reference's; its cost is O(events + observations)
// instead of O(sub-steps).  Observations happen at the end of sub-step
// it * n_sub, after that sub-step's scattering (radiativeTransfer.m:72-77).
//
// Event-driven process (material.timeSteps = 1, propagateParticle.m): the
// free flight between scatterings is exponential with mean mft(mode).
//
// A launch holds one or more problems (material + source + particle count +
// outputs) sharing the geometry, the observation bins and the time axis.  A
// work-group works on one problem at a time: it picks the problem with the
// most estimated work left per work-group already on it, its wavefronts
// draw particles of that problem from the problem's counter until none are
// left, and it then picks again.  With one problem this is a plain
// persistent kernel with dynamic particle distribution.
//
// Histogram placement (forward), template parameter HM:
//  HM_LDS    LDS-privatised per work-group when the whole histogram of a
//            problem fits in LDS (flushed with global atomics when the
//            work-group leaves the problem);
//  HM_GLOBAL global no-return atomics into E;
//  HM_TMAJOR global atomics into a time-major scratch copy T (time fastest,
//            the time axis padded to even length) that is added into E by a
//            tiled transpose afterwards.  A particle often stays in the same
//            bin over consecutive observation times; a deposit at an even time
//            waits one observation for its partner and the two go out as one
//            64-bit atomic on the aligned pair of counters (no carry can cross:
//            a counter never exceeds 2^32 - 1).  The histogram atomics are
//            executed at the memory side at a fixed request rate, so fewer
//            requests is what makes the kernel faster.  Counters hit by many
//            particles (near the source at early times every particle lands in
//            the same few bins, and same-address atomics serialise) are
//            aggregated in a small per-work-group LDS cache first: a
//            direct-mapped table of (T index, count) slots, each claimed by the
//            first index that maps to it and flushed when the work-group leaves
//            the problem.
//            Runs of three (LaunchParams::runs3): the scratch holds 64-bit
//            words of three 21-bit counters for three consecutive times, in
//            three copies shifted by one time, so that a run of up to three
//            consecutive deposits into one bin is one atomic whatever time it
//            starts at.  A 21-bit counter cannot overflow: particle id goes to
//            scratch copy id / run_n (run_n <= 2^21 - 1), a particle adds at
//            most one count to a counter, and the hot-counter cache, whose
//            counts are larger, is flushed straight into E.
//
// Adjoint: the same kernel replays every trajectory from its Philox counter
// (same draws, same events, same deposits) and, at each deposit into bin b
// at observation time t, adds weight[b] * S(t) to the gradient, where S(t)
// is the score of the trajectory up to t with respect to (c, lc, Q):
//  - scattering test of each sub-step, probability 1 - exp(-lam),
//    lam = dt_s Sigma: after k sub-steps with n scatterings
//    d log L / d log lam = n lam / expm1(lam) - (k - n) lam  (= A);
//    Sigma scales as c^2 (every coefficient of variation proportional to c,
//    MaterialClass.m:233-235), so dlam/dc = 2 lam / c and the angle law does
//    not depend on c;
//  - scattering angle theta with density sigma(theta) J(theta) / Sigma:
//    d log f / d lc = dlogsigma_dlc(theta) - dlogSigma_dlc, plus A dlogSigma_dlc
//    from the rate;
//  - Q removal test of each sub-step, survival exp(-omega dt_s / Q): after k
//    sub-steps alive, d log L / dQ = k omega dt_s / Q^2 (the pathwise
//    derivative of the survival weight, propagateParticleSmallDt.m:64-76).
// Event-driven: A = n - t / mft (Poisson process of rate 1/mft), no Q.
#pragma once

#include "rtmc_common.hpp"

namespace rtmc {

constexpr int64_t kInf = (int64_t)1 << 60;

// One problem of a launch, in device memory.
struct ProbDev {
  MatDev M;
  rtmc_source S;
  int64_t n;          // particles
  uint64_t pid0;      // global id of particle 0
  uint32_t k0, k1;    // Philox key
  int64_t nsub;       // sub-steps per observation interval
  double dt_sub;
  double lamS[2];     // scattering: dt_s / mft(mode) (per sub-step)
  double lamK[2];     // removal: omega dt_s / Q(mode) (per sub-step)
  double ilamS[2];    // 1 / lamS (0 where lamS is 0)
  double ilamK[2];    // 1 / lamK (0 where lamK is 0)
  int useQ;
  double cost;        // estimated work per particle (scheduling weight)
  uint32_t* E;        // histogram (forward)
  uint32_t* T;        // HM_TMAJOR scratch: T[(spatial + mode * time_stride) * nt_pad + it]
  uint32_t* T2;       // the same shifted by one time (T2[... + it - 1]) for pairs that start at odd times
                      // runs of three: T holds copies c = 0, 1, ... (run_words 64-bit words
                      // each) of U[(row * 3 + s) * tw + w], row = spatial + mode * tsp, whose
                      // word covers times 3w + s + {0, 1, 2} in bits 0-20, 21-41, 42-62
  unsigned long long* counters;  // RTMC_NCOUNTERS or null
  // adjoint
  const float* w;     // dL/dE per bin
  double* grad;       // [3]
  double* dhist;      // [3][hist] or null
  int64_t dhist_stride;
  double two_over_c;  // d log lam / dc
  double lam_over_expm1;  // lamS[0] / expm1(lamS[0])
  double dlogS;       // d log Sigma / d lc
  double kq;          // omega dt_s / Q^2 (0 when Q is infinite)
  const double* glc;  // d log sigma(theta) / d lc at theta_k = k pi / (n_glc - 1)
  int64_t n_glc;
  // guide of each inverse-CDF table in global memory (LaunchParams::gg_log2 > 0):
  // table t's cell c (u in [c/G, (c+1)/G)) at gg[t * (G + 1) + c] holds the largest
  // knot index k <= n-2 with x[k] <= c/G (0 below x[0])
  const int* gg;
};

// HM_TMAJOR3: HM_TMAJOR with runs of three (LaunchParams::runs3), its own instances
enum { HM_GLOBAL = 0, HM_LDS = 1, HM_TMAJOR = 2, HM_TMAJOR3 = 3 };

struct LaunchParams {
  rtmc_geometry G;
  const double* edges[4];
  int ne[4];
  int eoff[4];        // offset of each edge array in the LDS copy
  uint32_t stride[4]; // bin strides (a bin index within one time slice is < time_stride < 2^31)
  unsigned present;   // variables that exist (bit v; z is absent in 2-D)
  int64_t time_stride;
  uint32_t time_stride32;  // the same, for 32 x 32 -> 64-bit products
  int64_t mode_stride;
  int n_times;
  int dirvar;         // RTMC_DIRVAR_*: the direction variable is dir_z, or the cosine to the position
  double dt_obs;
  int edges_total;    // doubles of the LDS edge copy (0: edges read from global memory)
  int chunk;          // particles fetched per counter access (per wavefront)
  int batch;          // lanes waiting for an event before a wavefront takes an event step
  int birth_min;      // lanes wanting a particle before an event step births them (while others can fly)
  int ntab;           // inverse-CDF tables (1 acoustic, 4 elastic)
  int guide_n;        // LDS guide cells per table (power of two; 0 = none)
  int gg_log2;        // log2 of the global guide's cells per table (0 = none; then the LDS guide)
  int64_t hist_size;  // words of the LDS histogram / weight table (0 = global memory)
  int nt_pad;         // HM_TMAJOR: n_times rounded up to even (time stride of T is 1)
  int hot_slots;      // HM_TMAJOR: slots of the LDS cache of hot counters (power of two, or 0)
  int hot_shift;      // 32 - log2(hot_slots)
  int hot_off;        // byte offset of the cache in the dynamic LDS block
  int hot_times;      // only deposits at observation indices below this go through the cache
  unsigned hot_max_taken;  // particles a work-group takes from a problem before it flushes the cache (< 2^24)
  // HM_TMAJOR, 2-D with dir_z binned outermost: every particle's direction variable is
  // acos(~0) = pi/2, so the scratch holds only that direction bin (BinSetup::tdir)
  int tcomp;
  uint32_t tsp;       // spatial cells per mode in the scratch (time_stride, or stride[3] when tcomp)
  int odd_pairs;      // pairs also start at odd times (into T2)
  int runs3;          // HM_TMAJOR: runs of up to three times per atomic (T holds 21-bit counters)
  int run64;          // runs of three: a run of one as a 64-bit atomic too (else 32-bit on the low half)
  int tw;             // runs of three: words per row and shift, ceil(n_times / 3)
  int run_n;          // runs of three: particles per scratch copy (<= 2^21 - 1)
  int64_t run_words;  // runs of three: 64-bit words of one scratch copy
  // compressed inverse-CDF tables held in LDS (lt_cap > 0; MatDev::lk): table t's
  // knot records at byte lt_off + 16 t lt_cap, its guide (2^lt_glog2 + 1 uint16:
  // cell c holds the largest knot k <= m-2 with u_k <= c / 2^lt_glog2, 0 below
  // u_0) at byte lt_goff + 2 t lt_gstride
  int lt_cap;
  int lt_glog2;
  int lt_gstride;
  int lt_off;
  int lt_goff;
  int n_prob;
  unsigned long long* work;  // per-problem particle counters (zeroed)
  unsigned int* active;      // per-problem count of work-groups on it (zeroed)
};

// Per-particle Philox stream: counter = {c, 0, pid_lo, pid_hi}, key = (k0, k1).
// Every draw site takes whole blocks of four 32-bit words, so all lanes of a
// wavefront run the generator at the same sites, once per site (a per-lane
// leftover would desynchronise the lanes, and the generator would then run at
// every draw for some of them).  A uniform takes 53 bits (two words) where its
// tail matters -- the exponential and geometric free-flight draws and the
// Box-Muller radius -- and 32 bits (one word) where it is compared with a
// probability or turned into an angle: azimuths, the polar angle's table
// lookup, mode conversion, boundary decisions, birth directions.  2^-32 is far
// below the 1e-8 cdf steps of the reference's angle tables.
struct Rng {
  uint32_t k0, k1, plo, phi;
  uint32_t c;
  __device__ __forceinline__ void start(uint64_t pid) {
    plo = (uint32_t)pid;
    phi = (uint32_t)(pid >> 32);
    c = 0u;
  }
  __device__ __forceinline__ U4 block() {
    U4 in;
    in.x = c;
    in.y = 0u;
    in.z = plo;
    in.w = phi;
    c += 1u;
    return philox10(in, k0, k1);
  }
  // uniform in (0, 1] from one word
  __device__ __forceinline__ static double w32(uint32_t w) { return ((double)w + 1.0) * (1.0 / 4294967296.0); }
  // a free-flight uniform of the flight precision R: 53 bits from words (a, b), or 32 from a
  template <typename R>
  __device__ __forceinline__ static double wflight(uint32_t a, uint32_t b) {
    return sizeof(R) == 8 ? u53(a, b) : w32(a);
  }
  // the uniforms of one scattering: azimuth, polar angle and conversion (32 bits)
  // and the next free flight (53 bits with FP64 flight): one block, or two for
  // an elastic medium with FP64 flight
  template <typename R, bool AC>
  __device__ __forceinline__ void scatter4(double& uphi, double& uth, double& uconv, double& ufl) {
    const U4 r = block();
    uphi = w32(r.x);
    uth = w32(r.y);
    if (sizeof(R) != 8) {
      uconv = w32(r.z);
      ufl = w32(r.w);
    } else if (AC) {
      uconv = 0.0;
      ufl = u53(r.z, r.w);
    } else {
      uconv = w32(r.z);
      const U4 r2 = block();
      ufl = u53(r2.x, r2.y);
    }
  }
  // two free-flight uniforms of the flight precision R (one block)
  template <typename R>
  __device__ __forceinline__ void flight2(double& a, double& b) {
    const U4 r = block();
    a = wflight<R>(r.x, r.y);
    b = sizeof(R) == 8 ? u53(r.z, r.w) : w32(r.y);
  }
  // two 32-bit uniforms (one block)
  __device__ __forceinline__ void pair32(double& a, double& b) {
    const U4 r = block();
    a = w32(r.x);
    b = w32(r.y);
  }
};

// Words of the generator taken one at a time (a birth's draws; every lane of a
// birth follows the same sequence, so the blocks stay aligned across lanes).
struct WordStream {
  Rng& r;
  U4 b;
  int n;
  __device__ __forceinline__ explicit WordStream(Rng& rng) : r(rng), n(0) {}
  __device__ __forceinline__ uint32_t word() {
    if (n == 0) {
      b = r.block();
      n = 4;
    }
    const uint32_t w = n == 4 ? b.x : (n == 3 ? b.y : (n == 2 ? b.z : b.w));
    n -= 1;
    return w;
  }
  __device__ __forceinline__ double u32() { return Rng::w32(word()); }
  __device__ __forceinline__ double u53() {
    const uint32_t a = word();
    return rtmc::u53(a, word());
  }
};

// index of the next sub-step with a success, trials of probability 1-exp(-lam)
// (ilam = 1 / lam), from the uniform u
template <typename R>
__device__ __forceinline__ int64_t geom_from(double u, double lam, double ilam) {
  if (!(lam > 0.0)) return kInf;
  double e = (sizeof(R) == 8) ? -log(u) : (double)(-logf((float)u));
  double g = ceil(e * ilam);
  if (g < 1.0) g = 1.0;
  if (g > 1e15) return kInf;
  return (int64_t)g;
}

// First sub-step (relative to the segment base) at which the straight
// segment x = xb + n h d may trigger boundary B; a prediction that is at most
// one sub-step early (the sub-step is then executed in full and found clear).
template <typename R>
__device__ __forceinline__ int64_t predict_boundary(const rtmc_boundary& B, const R xb_[3], const R d_[3], R h_) {
  const double xb[3] = {xb_[0], xb_[1], xb_[2]};
  const double d[3] = {d_[0], d_[1], d_[2]};
  const double h = h_;
  if (B.dir < 4) {
    int b = B.dir - 1;
    double s0 = comp3(xb, b) - B.val;
    double r = h * comp3(d, b);
    if (!(s0 * r < 0.0)) return kInf;  // moving away, parallel, or on the plane
    double ns = fabs(s0) / fabs(r);
    if (!(ns < 1e15)) return kInf;
    int64_t nstar = (int64_t)floor(ns) + 1;
    return nstar > 1 ? nstar - 1 : 1;
  }
  double ax = xb[0], ay = xb[1];
  double bx = h * d[0], by = h * d[1];
  double a2 = ax * ax + ay * ay - B.val * B.val;
  if (a2 > 0.0) return 1;
  double bb = bx * bx + by * by;
  if (!(bb > 0.0)) return kInf;
  double ab = ax * bx + ay * by;
  double disc = ab * ab - bb * a2;
  if (disc < 0.0) disc = 0.0;
  double np = (-ab + sqrt(disc)) / bb;
  if (!(np < 1e15)) return kInf;
  int64_t f = (int64_t)floor(np);
  return f > 1 ? f : 1;
}

// Binning of the observation variables, set up once per work-group from the
// edges: which variables need a range check / bin search, and the mean
// spacing used for the first guess of the bin.
struct BinSetup {
  unsigned check;  // variables to bin / range-check (bit v)
  double e0[4];
  double ehi[4];   // last edge
  double invw[4];
  // radius the only binned position variable, edges >= 0 and finite: the bin is
  // found from the squared radius against squared edges (held in LDS in place of
  // the edges), with an FP32 square root for the first guess
  unsigned radial;
  float e0f, invwf;
  double s_lo, s_hi;  // squared first and last edges
  int tdir;           // LaunchParams::tcomp: the direction bin of acos(cos(pi/2)) (the scratch's only one)
};

// Is a one-bin variable's range check always passed?  (then it is skipped)
__device__ __forceinline__ bool trivially_inside(int v, int dim, int frame, double lo, double hi) {
  if (v == 3) {  // direction: dir_z clamped to [-1,1] (3-D) or acos of it in [0, pi] (2-D)
    return dim == 3 ? (lo <= -1.0 && hi >= 1.0) : (lo <= 0.0 && hi >= kPi);
  }
  if (frame == RTMC_FRAME_CARTESIAN || (v == 2 && frame == RTMC_FRAME_CYLINDRICAL)) {
    return lo == -INFINITY && hi == INFINITY;  // a Cartesian coordinate
  }
  if (v == 0) return lo <= 0.0 && hi == INFINITY;  // radius
  if (v == 1) return lo <= -kPi && hi >= kPi;      // atan2 azimuth
  return lo <= -1.0 && hi >= 1.0;                  // sin(elevation)
}

__device__ __forceinline__ void bin_setup(const LaunchParams& p, int dim, const double* sedges, BinSetup* bs) {
  bs->check = 0u;
  for (int v = 0; v < 4; ++v) {
    bs->e0[v] = 0.0;
    bs->ehi[v] = 0.0;
    bs->invw[v] = 0.0;
    if (!(p.present & (1u << v))) continue;
    const double* e = sedges ? sedges + p.eoff[v] : p.edges[v];
    double lo = e[0], hi = e[p.ne[v] - 1];
    // (the position-angle direction variable is checked: it is NaN at the origin)
    if (p.ne[v] == 2 && !(v == 3 && p.dirvar == RTMC_DIRVAR_POSITION) && trivially_inside(v, dim, p.G.frame, lo, hi))
      continue;
    bs->check |= 1u << v;
    bs->e0[v] = lo;
    bs->ehi[v] = hi;
    double w = hi - lo;
    bs->invw[v] = (isfinite(w) && w > 0.0) ? (double)(p.ne[v] - 1) / w : 0.0;
  }
  const bool polar = dim == 2 ? p.G.frame != RTMC_FRAME_CARTESIAN
                              : (p.G.frame == RTMC_FRAME_SPHERICAL || p.G.frame == RTMC_FRAME_CYLINDRICAL);
  bs->radial = (sedges != nullptr && polar && (bs->check & 7u) == 1u && bs->e0[0] >= 0.0 && isfinite(bs->ehi[0]) &&
                bs->invw[0] > 0.0)
                   ? 1u
                   : 0u;
  bs->e0f = (float)bs->e0[0];
  bs->invwf = (float)bs->invw[0];
  bs->s_lo = bs->e0[0] * bs->e0[0];
  bs->s_hi = bs->ehi[0] * bs->ehi[0];
  bs->tdir = -1;
  if (p.tcomp) {
    const double* e3 = sedges ? sedges + p.eoff[3] : p.edges[3];
    bs->tdir = (int)bin_search(e3, p.ne[3], acos(fmin(1.0, fmax(-1.0, cos(kPi / 2)))));
  }
}

// Birth (initializeParticle.m:17-147), computed in FP64 for both flight
// precisions (once per particle; keeps dir_z = cos(pi/2) of 2-D births at its
// FP64 value).
template <int DIM>
__device__ __forceinline__ void birth64(const rtmc_source& S, Rng& rng, double x[3], double d[3], double q[3]) {
  // The same variates and the same angles as initializeParticle.m:17-147; each
  // angle 2 pi u enters through sincospi(2u), and a polar angle acos(c) through
  // its cosine c and sine sqrt(1 - c^2).
  // Both branches fill the same locals and the arrays are written once at the
  // end: stores to different arrays at the ends of the two branches would be
  // merged into one store through a pointer select, which keeps the arrays in
  // the private segment.
  WordStream ws(rng);
  double x0, x1, x2, d0, d1, d2, q0, q1, q2;
  if (S.type == RTMC_SRC_POINT) {
    double r;
    if (S.radial.n == 0) {
      const double u1 = ws.u53(), u2 = ws.u32();
      double g = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
      r = fabs(g * S.lambda);  // initializeParticle.m:23
    } else {
      r = table_eval(gl(S.radial.x), gl(S.radial.y), S.radial.n, S.radial.extrap, ws.u53());  // :25-27
    }
    // polar angles as (cos, sin); azimuths as turns
    double cp0, sp0, cp, sp;
    if (DIM == 2) {
      cp0 = cp = cos(kPi / 2);  // 6.12e-17, as MATLAB's cos(pi/2)
      sp0 = sp = 1.0;
    } else {
      cp0 = -1.0 + 2.0 * ws.u32();
      sp0 = sqrt(fmax(0.0, 1.0 - cp0 * cp0));
      cp = -1.0 + 2.0 * ws.u32();
      sp = sqrt(fmax(0.0, 1.0 - cp * cp));
    }
    const double t0 = ws.u32();
    double t = ws.u32();
    if (S.direction == RTMC_DIR_OUTGOING) {
      t = t0;
      cp = cp0;
      sp = sp0;
    } else if (S.direction == RTMC_DIR_UPPER) {
      t = ws.u32();
      cp = ws.u32();
      sp = sqrt(fmax(0.0, 1.0 - cp * cp));
    }
    double st, ct;
    sincospi(2.0 * t, &st, &ct);
    d0 = ct * sp;
    d1 = st * sp;
    d2 = cp;
    q0 = -st;
    q1 = ct;
    q2 = 0.0;
    if (DIM == 3) {
      const double dv[3] = {d0, d1, d2}, qv[3] = {q0, q1, q2};
      double c[3];
      cross3(dv, qv, c);
      double sa, ca;
      sincospi(2.0 * ws.u32(), &sa, &ca);
      q0 = q0 * ca + c[0] * sa;
      q1 = q1 * ca + c[1] * sa;
      q2 = q2 * ca + c[2] * sa;
    }
    double st0, ct0;
    sincospi(2.0 * t0, &st0, &ct0);
    x0 = S.position[0] + r * (ct0 * sp0);
    x1 = S.position[1] + r * (st0 * sp0);
    x2 = S.position[2] + r * cp0;
  } else {
    const double u1 = ws.u53(), u2 = ws.u32();
    double g = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    double c_long = g * S.lambda / 2.0;
    double c1, c2;
    const double a = ws.u32(), b = ws.u32();
    if (S.aperture == RTMC_APERTURE_RECTANGLE) {
      c1 = (a - 0.5) * S.extent[0];
      c2 = (b - 0.5) * S.extent[1];
    } else {
      double rho = S.extent[0] * sqrt(a);
      double sa, ca;
      sincospi(2.0 * b, &sa, &ca);
      c1 = rho * ca;
      c2 = rho * sa;
    }
    x0 = ((S.position[0] + c_long * S.normal[0]) + c1 * S.basis_u[0]) + c2 * S.basis_v[0];
    x1 = ((S.position[1] + c_long * S.normal[1]) + c1 * S.basis_u[1]) + c2 * S.basis_v[1];
    x2 = ((S.position[2] + c_long * S.normal[2]) + c1 * S.basis_u[2]) + c2 * S.basis_v[2];
    d0 = S.normal[0];
    d1 = S.normal[1];
    d2 = S.normal[2];
    double sn, cs;
    sincospi(2.0 * ws.u32(), &sn, &cs);
    q0 = cs * S.basis_u[0] + sn * S.basis_v[0];
    q1 = cs * S.basis_u[1] + sn * S.basis_v[1];
    q2 = cs * S.basis_u[2] + sn * S.basis_v[2];
  }
  x[0] = x0;
  x[1] = x1;
  x[2] = x2;
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
  q[0] = q0;
  q[1] = q1;
  q[2] = q2;
}

template <int DIM, typename R>
__device__ __forceinline__ void birth(const rtmc_source& S, Rng& rng, R xo[3], R dd[3], R qo[3]) {
  double x[3], d[3], q[3];
  birth64<DIM>(S, rng, x, d, q);
  for (int k = 0; k < 3; ++k) {
    xo[k] = (R)x[k];
    dd[k] = (R)d[k];
    qo[k] = (R)q[k];
  }
}

// Bin of the direction variable (observeTime.m:24-32: dir_z in 3-D, acos(dir_z)
// in 2-D, as obs_vars computes it).  It changes only with the direction, so a
// lane keeps it from one event to the next.  0 when the variable needs no
// check, -1 when outside its edges.
template <typename R>
__device__ __forceinline__ int dir_bin(const LaunchParams& p, const BinSetup& bs, const double* __restrict__ sedges,
                                       int dim, R dz) {
  if (!(bs.check & 8u)) return 0;
  R cp = dz;
  if (cp > R(1)) cp = R(1);
  if (cp < R(-1)) cp = R(-1);
  const R psi = (dim == 3) ? cp : acos(cp);
  // two calls rather than one through a selected pointer: each reads its own
  // memory (LDS or global) instead of through flat addressing
  if (p.edges_total > 0) return bin_guess2(sedges + p.eoff[3], p.ne[3], bs.e0[3], bs.ehi[3], bs.invw[3], (double)psi);
  return bin_guess2(p.edges[3], p.ne[3], bs.e0[3], bs.ehi[3], bs.invw[3], (double)psi);
}

// Index of the observation's bin over the binned variables (observeTime.m:15-45
// + histcounts), without the time and mode terms, or -1 when outside; the
// direction's bin `dbin` comes from dir_bin.
template <typename R, bool POS>
__device__ __forceinline__ int bin_index(const LaunchParams& p, const BinSetup& bs,
                                         const double* __restrict__ sedges, int dim, const R X[3], int dbin,
                                         const R d[3]) {
  uint32_t idx = 0;
  if (bs.check & 8u) {
    // the direction's bin comes with the direction (dir_bin), unless the variable
    // is the cosine to the position vector, which moves with the particle
    if (POS) dbin = dir_bin(p, bs, sedges, dim, cos_to_position(d, X));
    if (dbin < 0) return -1;
    idx = (uint32_t)dbin * p.stride[3];
  }
  if (bs.radial) {
    // e_k <= r < e_k+1  <=>  e_k^2 <= r^2 < e_k+1^2 for edges >= 0 (last bin closed)
    const double x = X[0], y = X[1], z = X[2];
    const double s = (dim == 3 && p.G.frame == RTMC_FRAME_SPHERICAL) ? x * x + y * y + z * z : x * x + y * y;
    if (!(s >= bs.s_lo && s <= bs.s_hi)) return -1;
    const int nb = p.ne[0] - 1;
    const float g = (__builtin_amdgcn_sqrtf((float)s) - bs.e0f) * bs.invwf;  // guess only: v_sqrt_f32 as is
    int k = (int)fminf(fmaxf(g, 0.0f), (float)(nb - 1));  // NaN -> 0
    const double* e2 = sedges + p.eoff[0];
    const double lo = e2[k], hi = e2[k + 1];
    if (!((s >= lo) & ((s < hi) | (k == nb - 1)))) {
      while (k > 0 && s < e2[k]) --k;
      while (k < nb - 1 && s >= e2[k + 1]) ++k;
    }
    return (int)(idx + (uint32_t)k * p.stride[0]);
  }
  double v[4];
  obs_vars<R, true>(dim, p.G.frame, X, R(0), bs.check & 7u, v);
  if (p.edges_total > 0) {
    // edges in LDS (`sedges` points into the work-group's LDS block)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!(bs.check & (1u << k))) continue;
      int b = bin_guess2(sedges + p.eoff[k], p.ne[k], bs.e0[k], bs.ehi[k], bs.invw[k], v[k]);
      if (b < 0) return -1;
      idx += (uint32_t)b * p.stride[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!(bs.check & (1u << k))) continue;
      int b = bin_guess2(p.edges[k], p.ne[k], bs.e0[k], bs.ehi[k], bs.invw[k], v[k]);
      if (b < 0) return -1;
      idx += (uint32_t)b * p.stride[k];
    }
  }
  return (int)idx;
}

// d log sigma / d lc at theta (linear interpolation on the uniform grid)
__device__ __forceinline__ double glc_eval(const ProbDev& P, double th) {
  const double h = kPi / (double)(P.n_glc - 1);
  double s = th / h;
  if (!(s > 0.0)) s = 0.0;
  const double smax = (double)(P.n_glc - 1);
  if (s > smax) s = smax;
  int64_t i = (int64_t)s;
  if (i > P.n_glc - 2) i = P.n_glc - 2;
  double f = s - (double)i;
  return P.glc[i] + f * (P.glc[i + 1] - P.glc[i]);
}

// Problem choice of a work-group: most estimated work left per work-group on
// it (+1 for the one choosing).  Returns -1 when every problem is handed out.
__device__ __forceinline__ int pick_problem(const LaunchParams& p, const ProbDev* __restrict__ probs) {
  int best = -1;
  double bestv = 0.0;
  for (int i = 0; i < p.n_prob; ++i) {
    unsigned long long w = __hip_atomic_load(p.work + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int64_t n = probs[i].n;
    if ((int64_t)w >= n) continue;
    unsigned a = __hip_atomic_load(p.active + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double v = (double)(n - (int64_t)w) * probs[i].cost / (double)(a + 1u);
    if (v > bestv) {
      bestv = v;
      best = i;
    }
  }
  if (best >= 0) atomicAdd(p.active + best, 1u);
  return best;
}

// SINK: 0 forward histogram, 1 adjoint (weighted scores)
// LDS-histogram kernels run one 1024-lane work-group per CU (the histogram
// takes most of the LDS); the others use 256-lane work-groups.
// POS: the direction variable is the cosine to the position vector (RTMC_DIRVAR_POSITION)
// LDS the tracing kernel declares statically, beside the dynamic block the plan
// sizes: the bin set-up, the angle-table ranges and two work-sharing words
// (with room for their alignment).
constexpr int kTraceStaticLds = 256;
static_assert(sizeof(BinSetup) + 8 * sizeof(double) + 2 * sizeof(int) + 32 <= (size_t)kTraceStaticLds,
              "kTraceStaticLds must cover the kernel's static LDS");

// LT: scattering angles from the compressed tables held in LDS (LaunchParams::lt_cap
// > 0; its own instances, so the full-table instances keep their registers)
template <int DIM, int HM, bool QUANT, bool AC, bool BND, typename R, int SINK, bool POS, bool LT>
__global__ __launch_bounds__((HM == HM_LDS || (DIM == 3 && HM == HM_TMAJOR && !LT) || (DIM == 2 && HM == HM_TMAJOR3 && !LT)) ? 1024 : 256) void k_trace(const LaunchParams p,
                                                                     const ProbDev* __restrict__ probs) {
  constexpr bool LDSH = HM == HM_LDS;
  constexpr bool TMAJ = HM == HM_TMAJOR || HM == HM_TMAJOR3;
  constexpr bool RUNS3 = HM == HM_TMAJOR3;
  extern __shared__ double smem[];
  double* sedges = smem;
  __shared__ BinSetup s_bs;
  __shared__ double s_tx0[4], s_tx1[4];
  __shared__ int s_cur;
  __shared__ unsigned int s_taken;  // particles this work-group took from the current problem
  int* sguide = (int*)(smem + p.edges_total);
  uint32_t* shist = (uint32_t*)(sguide + (p.guide_n > 0 ? p.ntab * (p.guide_n + 1) : 0));
  float* sw = (float*)shist;  // adjoint: the weight table takes the histogram's place
  constexpr bool ADJ = SINK == 1;
  // HM_TMAJOR cache of hot counters: one 64-bit LDS word per slot holding the T index
  // (bits 24-63) and its count (bits 0-23; a counter gets one deposit per particle
  // and the host keeps a work-group's particles below 2^24)
  constexpr unsigned long long kEmpty = ~0ull;
  unsigned long long* hslot = (unsigned long long*)((char*)smem + p.hot_off);
  rtmc_lds_knot* sknots = (rtmc_lds_knot*)((char*)smem + p.lt_off);
  uint16_t* sltg = (uint16_t*)((char*)smem + p.lt_goff);

  if (p.edges_total > 0) {
    // gather the edge arrays into one LDS block
    for (int v = 0; v < 4; ++v) {
      if (!(p.present & (1u << v))) continue;
      for (int k = threadIdx.x; k < p.ne[v]; k += blockDim.x) smem[p.eoff[v] + k] = p.edges[v][k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) bin_setup(p, DIM, p.edges_total > 0 ? sedges : nullptr, &s_bs);
  __syncthreads();
  if (s_bs.radial) {
    // squared radial edges for bin_index
    for (int k = threadIdx.x; k < p.ne[0]; k += blockDim.x) {
      const double e = smem[p.eoff[0] + k];
      smem[p.eoff[0] + k] = e * e;
    }
    __syncthreads();
  }
  const BinSetup& bs = s_bs;  // read from LDS where used (wave-uniform values, no registers held)
  const int lane = threadIdx.x & 63;

  while (true) {
    if (threadIdx.x == 0) {
      s_cur = pick_problem(p, probs);
      s_taken = 0u;
    }
    __syncthreads();
    const int cur = __builtin_amdgcn_readfirstlane(s_cur);
    if (cur < 0) break;
    const ProbDev& P = probs[cur];
    RTMC_GLOBAL uint32_t* const PE = gl(P.E);
    RTMC_GLOBAL uint32_t* const PT = gl(P.T);
    const MatDev& M = P.M;
    unsigned long long* const work = p.work + cur;

    if (p.guide_n > 0) {
      // guide table: cell c holds the largest knot index k <= n-2 with x[k] <= c/G
      // (Chen & Asau's guide for inversion of a tabulated CDF)
      const int G = p.guide_n;
      for (int t = 0; t < p.ntab; ++t) {
        const TableDev& T = M.inv[t];
        if (threadIdx.x == 0) {
          s_tx0[t] = T.x[0];
          s_tx1[t] = T.x[T.n - 1];
        }
        for (int c = threadIdx.x; c <= G; c += blockDim.x) {
          double q = (double)c / (double)G;
          int64_t lo = 0, hi = T.n - 1;
          if (q >= T.x[0]) {
            while (hi - lo > 1) {
              int64_t mid = lo + ((hi - lo) >> 1);
              if (T.x[mid] <= q) lo = mid;
              else hi = mid;
            }
          }
          sguide[t * (G + 1) + c] = (int)lo;
        }
      }
    }
    if (LT) {
      // the problem's compressed tables into LDS, then their guides; the clamp
      // bounds s_tx0 / s_tx1 hold NEAREST lookups to the end knots
      for (int t = 0; t < p.ntab; ++t) {
        const int m = M.lm[t];
        const RTMC_GLOBAL rtmc_lds_knot* src = gl(M.lk[t]);
        rtmc_lds_knot* dst = sknots + t * p.lt_cap;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
          dst[k].u = src[k].u;
          dst[k].theta = src[k].theta;
          dst[k].slope = src[k].slope;
        }
        if (threadIdx.x == 0) {
          s_tx0[t] = M.lnear ? src[0].u : -INFINITY;
          s_tx1[t] = M.lnear ? src[m - 1].u : INFINITY;
        }
      }
      __syncthreads();
      const int G = 1 << p.lt_glog2;
      const double invG = 1.0 / (double)G;  // exact
      for (int t = 0; t < p.ntab; ++t) {
        const int m = M.lm[t];
        const rtmc_lds_knot* kn = sknots + t * p.lt_cap;
        for (int c = threadIdx.x; c <= G; c += blockDim.x) {
          const double q = (double)c * invG;
          int lo = 0, hi = m - 1;
          if (q >= kn[0].u) {
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (kn[mid].u <= q) lo = mid;
              else hi = mid;
            }
          }
          sltg[t * p.lt_gstride + c] = (uint16_t)lo;
        }
      }
    }
    if (LDSH) {
      if (ADJ) {
        for (int64_t k = threadIdx.x; k < p.hist_size; k += blockDim.x) sw[k] = P.w[k];
      } else {
        for (int64_t k = threadIdx.x; k < p.hist_size; k += blockDim.x) shist[k] = 0u;
      }
    }
    if (TMAJ && !ADJ) {
      for (int k = threadIdx.x; k < p.hot_slots; k += blockDim.x) hslot[k] = kEmpty;
    }
    __syncthreads();

    // inverse-CDF lookup of the reference table (same interval and formula as
    // table_eval), bracketed by the LDS guide when present
    auto lookup = [&](int t, double u) -> double {
      if (LT) {
        // compressed table in LDS: the guide brackets the knot, a short search in
        // LDS finds it, and the chord from it gives theta
        const double q = fmin(fmax(u, s_tx0[t]), s_tx1[t]);
        const int G = 1 << p.lt_glog2;
        int c = (int)(q * (double)G);
        c = c < 0 ? 0 : (c > G - 1 ? G - 1 : c);
        const uint16_t* g = sltg + t * p.lt_gstride;
        int lo = g[c], hi = (int)g[c + 1] + 1;
        const rtmc_lds_knot* kn = sknots + t * p.lt_cap;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (kn[mid].u <= q) lo = mid;
          else hi = mid;
        }
        return (double)kn[lo].theta + (q - kn[lo].u) * (double)kn[lo].slope;
      }
      const TableDev& T = M.inv[t];
      const RTMC_GLOBAL double* const Tx = gl(T.x);
      const RTMC_GLOBAL double* const Ty = gl(T.y);
      if (p.gg_log2 > 0) {
        // cell of u in [0, 1] (u * G is exact for G a power of two); one 8-byte read
        // brackets the knot, usually to one interval
        const int G = 1 << p.gg_log2;
        if (u >= 0.0 && u <= 1.0) {
          int c = (int)(u * (double)G);
          c = c > G - 1 ? G - 1 : c;
          const RTMC_GLOBAL int* g = gl(P.gg) + (size_t)t * (size_t)(G + 1) + c;
          return table_eval_exact_bracket(Tx, Ty, T.n, T.extrap, u, g[0], (int64_t)g[1] + 1);
        }
        return table_eval(Tx, Ty, T.n, T.extrap, u);
      }
      if (p.guide_n > 0 && u >= s_tx0[t] && u <= s_tx1[t]) {
        const int G = p.guide_n;
        int c = (int)(u * (double)G);
        c = c < 0 ? 0 : (c > G - 1 ? G - 1 : c);
        const int* g = sguide + t * (G + 1);
        return table_eval_bracket(Tx, Ty, T.n, T.extrap, u, g[c], (int64_t)g[c + 1] + 1);
      }
      return table_eval(Tx, Ty, T.n, T.extrap, u);
    };
    uint32_t nscat = 0, ndep = 0, nbnd = 0, nkill = 0;  // this lane's, for this problem
    // adjoint sums of this lane: w A, w (sum over events of dlogsigma - dlogSigma), w k
    double gaA = 0.0, gaG = 0.0, gaK = 0.0;
    // ---- lane state machine ------------------------------------------------
    // Each lane holds one particle and its own next observation index.  One
    // iteration is either a DEPOSIT step (every flying lane deposits its next
    // observation) or an EVENT step (lanes whose next event comes before that
    // observation process it, idle lanes take a new particle).  The wavefront
    // takes an event step only when at least `batch` lanes wait for one (or no
    // lane can fly), so the expensive event path runs for many lanes at once
    // instead of once per observation for whichever lane happens to need it.
    Rng rng;
    rng.k0 = P.k0;
    rng.k1 = P.k1;
    rng.start(0);
    R xb[3] = {R(0), R(0), R(0)}, d[3] = {R(0), R(0), R(1)}, q[3] = {R(1), R(0), R(0)};
    int isP = 1;
    R h = R(0);                     // sub-stepped: step length v dt_s; event-driven: speed v
    int64_t kb = 0, Ks = kInf, Kk = kInf, kev = kInf;   // sub-stepped: base, next scatter / removal / event
    double tb = 0.0, ts = INFINITY;                     // event-driven: base time, next scattering time
    int it = 1;                     // next observation index
    bool has = false;               // lane holds a particle
    bool need = true;               // lane wants a particle
    int64_t tpend = -1;             // HM_TMAJOR: T index of a deposit waiting for its pair
    // HM_TMAJOR runs of three: the open run's scratch row (-1: none), first time,
    // length and scratch copy; the scratch copy of the lane's particle
    int jrow = -1, trun = 0, lrun = 0, rsub = 0, psub = 0;
    // one atomic for the open run: the word shifted to its first time
    auto flush_run = [&]() {
      const unsigned tr = (unsigned)trun;
      const unsigned w = tr / 3u, s = tr - 3u * w;
      RTMC_GLOBAL unsigned long long* a = (RTMC_GLOBAL unsigned long long*)PT + (size_t)(uint32_t)rsub * (size_t)p.run_words +
                              ((size_t)(uint32_t)jrow * 3u + s) * (size_t)(uint32_t)p.tw + w;
      if (lrun == 1 && !p.run64) {
        // the first counter (bits 0-20) alone: a 32-bit atomic on the low half, which
        // the memory side executes faster than a 64-bit one.  No carry leaves bit 20
        // (a copy serves fewer than 2^21 particles), and the device's atomics on the
        // two halves of one aligned 64-bit word act on the same memory as 64-bit
        // atomics on it, so mixing the two widths adds exactly.
        __hip_atomic_fetch_add((RTMC_GLOBAL uint32_t*)a, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_fetch_add(a, lrun == 1 ? 1ull : (lrun == 2 ? 0x200001ull : 0x40000200001ull), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
      jrow = -1;
    };
    int dbin = 0;                   // bin of the direction variable for the current direction
    int nsc = 0;                    // adjoint: scatterings of the particle so far
    double gsum = 0.0;              // adjoint: sum of dlogsigma_dlc(theta) - dlogSigma_dlc over them
    // adjoint: sum of w and of w * it over the deposits since the particle's last
    // scattering, whose scores share nsc and gsum and are linear in it
    double aw = 0.0, awi = 0.0;
    // adds those deposits' w A, w gsum and w k to the lane's sums (before nsc or
    // gsum change, and when the lane is done): sub-stepped A = nsc (lam/expm1(lam)
    // + lam) - it nsub lam and k = it nsub, event-driven A = nsc - it dt / mft
    auto fold = [&]() {
      if (!ADJ) return;
      if (QUANT) {
        gaA += (double)nsc * (P.lam_over_expm1 + P.lamS[0]) * aw - (double)P.nsub * P.lamS[0] * awi;
        gaK += (double)P.nsub * awi;
      } else {
        gaA += (double)nsc * aw - (p.dt_obs / M.mftP) * awi;
      }
      gaG += gsum * aw;
      aw = 0.0;
      awi = 0.0;
    };
    // wave-uniform pool of particle indices fetched from the problem's counter
    int64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;

    auto next_event = [&]() {
      // sub-stepped: next sub-step at which something happens to this particle
      int64_t kbd = kInf;
      if (BND) {
        for (int ib = 0; ib < p.G.n_bnd; ++ib) {
          int64_t nb = predict_boundary(p.G.bnd[ib], xb, d, h);
          if (nb < kInf) kbd = min(kbd, kb + nb);
        }
      }
      kev = min(Ks, min(Kk, kbd));
    };

    while (true) {
      const bool pend_ev = has && (QUANT ? (kev <= (int64_t)it * P.nsub) : (ts <= (double)it * p.dt_obs));
      const bool pend = need || pend_ev;
      const bool fly = has && !pend_ev;
      const uint64_t pm = __ballot(pend);
      const uint64_t fm = __ballot(fly);
      if (pm == 0ull && fm == 0ull) break;
      if (__popcll(pm) >= p.batch || fm == 0ull) {
        // ---------------- EVENT step ----------------
        // Births wait until birth_min lanes want a particle, unless no lane can fly or
        // no event is pending: the birth path then runs for several lanes at once
        // instead of for the one or two that ran out in each event step.
        uint64_t nm = __ballot(need);
        if (nm != 0ull && __popcll(nm) < p.birth_min && fm != 0ull && (pm & ~nm) != 0ull) nm = 0ull;
        if (nm) {
          // hand out particle indices to the lanes that need one
          const int n = __popcll(nm);
          int64_t L = pool_end - pool_next;
          int64_t nb0 = 0, nb1 = 0;
          if (L < n && !exhausted) {
            unsigned long long base = 0;
            const int leader = __ffsll((long long)nm) - 1;
            if (lane == leader) {
              // a hot-counter slot counts at most one deposit per particle in 24 bits:
              // past kMaxTaken particles the work-group flushes its cache and picks again
              if (TMAJ && !ADJ && p.hot_slots > 0 && atomicAdd(&s_taken, (unsigned)p.chunk) >= p.hot_max_taken)
                base = (unsigned long long)P.n;
              else
                base = atomicAdd(work, (unsigned long long)p.chunk);
            }
            base = __shfl(base, leader, 64);
            if ((int64_t)base >= P.n) {
              exhausted = true;
            } else {
              nb0 = (int64_t)base;
              nb1 = min((int64_t)base + (int64_t)p.chunk, P.n);
            }
          }
          if (need) {  // nm != 0: this step births
            // rank among the birthing lanes: set bits of nm below this lane (mbcnt
            // counts them without a per-lane mask held in two registers)
            const int r = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(nm >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
            int64_t id = -1;
            if (r < L) id = pool_next + r;
            else if (r - L < nb1 - nb0) id = nb0 + (r - L);
            if (id >= 0) {
              rng.start(P.pid0 + (uint64_t)id);
              if (RUNS3) psub = (int)(id / p.run_n);
              birth<DIM>(P.S, rng, xb, d, q);
              double gs, gk;
              rng.flight2<R>(gs, gk);
              dbin = dir_bin(p, bs, sedges, DIM, d[2]);
              isP = (!AC && P.S.polarization == 1) ? 0 : 1;
              it = 1;
              has = true;
              need = false;
              fold();  // the lane's previous particle
              nsc = 0;
              gsum = 0.0;
              if (QUANT) {
                h = (R)((isP ? M.vp : M.vs) * P.dt_sub);
                kb = 0;
                Ks = geom_from<R>(gs, P.lamS[isP ? 0 : 1], P.ilamS[isP ? 0 : 1]);
                Kk = P.useQ ? geom_from<R>(gk, P.lamK[isP ? 0 : 1], P.ilamK[isP ? 0 : 1]) : kInf;
                next_event();
              } else {
                h = (R)(isP ? M.vp : M.vs);
                tb = 0.0;
                ts = -log(gs) * (isP ? M.mftP : M.mftS);
              }
            } else if (exhausted || r - L >= nb1 - nb0) {
              need = false;  // nothing left to trace
            }
          }
          if (n <= L) {
            pool_next += n;
          } else if (nb1 > nb0) {
            pool_next = nb0 + min((int64_t)(n - L), nb1 - nb0);
            pool_end = nb1;
          } else {
            pool_next = pool_end;
          }
        }
        if (pend_ev) {
          if (QUANT) {
            // sub-step j = kev in full (propagateParticleSmallDt.m:85-642); straight
            // flight from the base to the end of sub-step j-1 first
            const int64_t j = kev;
            const R m = (R)(j - 1 - kb);
            R x[3];
            for (int k = 0; k < 3; ++k) x[k] = xb[k] + (m * h) * d[k];
            bool alive = true;
            if (Kk == j) {  // removed at the start of sub-step j (:100-138)
              alive = false;
              nkill += 1;
            } else {
              R x1[3] = {x[0], x[1], x[2]};
              for (int k = 0; k < 3; ++k) x[k] = x[k] + h * d[k];
              const int wasP = isP;
              for (int ib = 0; BND && ib < p.G.n_bnd; ++ib) {
                const rtmc_boundary& B = p.G.bnd[ib];
                bool hit = (B.dir < 4) ? plane_crossed(B, x, x1) : cyl_outside(B, x);
                if (!hit) continue;
                nbnd += 1;
                if (B.type == RTMC_BND_ABSORBING) {
                  alive = false;
                  nkill += 1;
                  break;
                }
                double uA = 0.0, uB = 0.0;
                if (!AC) rng.pair32(uA, uB);
                if (B.dir < 4) plane_reflect(M, AC, B, x, d, q, isP, uA, uB);
                else cyl_reflect(M, AC, B, x, d, q, isP, uA, uB);
              }
              if (alive) {
                if (!AC && isP != wasP) {
                  // mode changed at the boundary: this sub-step's scattering test and
                  // the removal tests from the next one use the new mode
                  h = (R)((isP ? M.vp : M.vs) * P.dt_sub);
                  double gs, gk;
                  rng.flight2<R>(gs, gk);
                  Ks = j - 1 + geom_from<R>(gs, P.lamS[isP ? 0 : 1], P.ilamS[isP ? 0 : 1]);
                  Kk = P.useQ ? j + geom_from<R>(gk, P.lamK[isP ? 0 : 1], P.ilamK[isP ? 0 : 1]) : kInf;
                }
                if (Ks == j) {
                  const int before = isP;
                  // one draw site: azimuth, polar angle, conversion, next scattering
                  double uphi, uth, uconv, ugs;
                  rng.scatter4<R, AC>(uphi, uth, uconv, ugs);
                  double th = scatter_angles_t<DIM, R, decltype(lookup), true>(M, AC, d, q, isP, uphi < 0.5, uphi, uth, uconv, lookup);
                  if (ADJ) {
                    fold();
                    nsc += 1;
                    gsum += glc_eval(P, th) - P.dlogS;
                  }
                  rotate_renorm_fast(d, q, (R)th);
                  nscat += 1;
                  if (!AC && isP != before) {
                    h = (R)((isP ? M.vp : M.vs) * P.dt_sub);
                    if (P.useQ) {
                      double gk, unused;
                      rng.flight2<R>(gk, unused);
                      Kk = j + geom_from<R>(gk, P.lamK[isP ? 0 : 1], P.ilamK[isP ? 0 : 1]);
                    }
                  }
                  Ks = j + geom_from<R>(ugs, P.lamS[isP ? 0 : 1], P.ilamS[isP ? 0 : 1]);
                }
                xb[0] = x[0];
                xb[1] = x[1];
                xb[2] = x[2];
                kb = j;
                next_event();
                dbin = dir_bin(p, bs, sedges, DIM, d[2]);
              }
            }
            if (!alive) {
              has = false;
              need = true;
              if (TMAJ && tpend >= 0) {
                __hip_atomic_fetch_add(PT + tpend, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tpend = -1;
              }
              if (RUNS3 && jrow >= 0) flush_run();
            }
          } else {
            // event-driven: scattering at time ts (propagateParticle.m:24-72)
            const R f = (R)((double)h * (ts - tb));
            for (int k = 0; k < 3; ++k) xb[k] = xb[k] + f * d[k];
            tb = ts;
            double uphi, uth, uconv, uexp;
            rng.scatter4<R, AC>(uphi, uth, uconv, uexp);
            double th = scatter_angles_t<DIM, R, decltype(lookup), true>(M, AC, d, q, isP, uphi < 0.5, uphi, uth, uconv, lookup);
            if (ADJ) {
              fold();
              nsc += 1;
              gsum += glc_eval(P, th) - P.dlogS;
            }
            rotate_renorm_fast(d, q, (R)th);
            dbin = dir_bin(p, bs, sedges, DIM, d[2]);
            nscat += 1;
            h = (R)(isP ? M.vp : M.vs);
            ts = tb - log(uexp) * (isP ? M.mftP : M.mftS);
          }
        }
      } else {
        // ---------------- DEPOSIT step ----------------
        // Up to DEPS observations per lane (each lane stops at its next event or
        // its last observation): the step's scheduling is shared by them.
        constexpr int DEPS = TMAJ ? 1 : 8;
#pragma unroll
        for (int rep = 0; rep < DEPS; ++rep) {
        const bool dep_now =
            rep == 0 ? fly : (has && (QUANT ? (kev > (int64_t)it * P.nsub) : (ts > (double)it * p.dt_obs)));
        if (dep_now) {
          R X[3];
          if (QUANT) {
            const R m = (R)((int64_t)it * P.nsub - kb);
            for (int k = 0; k < 3; ++k) X[k] = xb[k] + (m * h) * d[k];
          } else {
            const R f = (R)((double)h * ((double)it * p.dt_obs - tb));
            for (int k = 0; k < 3; ++k) X[k] = xb[k] + f * d[k];
          }
          const int sp = bin_index<R, POS>(p, bs, sedges, DIM, X, dbin, d);
          const int64_t idx = sp < 0 ? -1
                                     : (int64_t)sp + (uint64_t)(uint32_t)it * p.time_stride32 +
                                           (isP ? 0 : p.mode_stride);
          if (TMAJ && !ADJ) {
            // time-major scratch: pair an even-time deposit with the next one.  With a
            // single-direction scratch (tcomp) a deposit in another direction bin, which
            // the 2-D process does not produce, goes straight into E.
            const bool tfit = !p.tcomp || dbin == bs.tdir;
            const uint32_t tbase = p.tcomp ? (uint32_t)bs.tdir * p.stride[3] : 0u;
            if (sp >= 0 && !tfit) {
              __hip_atomic_fetch_add(PE + idx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ndep += 1;
            }
            const int64_t j = (sp < 0 || !tfit)
                                  ? -1
                                  : (int64_t)((uint64_t)((uint32_t)sp - tbase + (isP ? 0u : p.tsp)) *
                                                  (uint32_t)p.nt_pad +
                                              (uint32_t)it);
            bool cached = false;
            if (j >= 0 && p.hot_slots > 0 && it < p.hot_times) {
              // multiplicative hash of the (< 2^40) T index folded to 32 bits
              const uint32_t hj = (uint32_t)((unsigned long long)j ^ ((unsigned long long)j >> 27));
              const int s = (int)((hj * 0x9E3779B1u) >> p.hot_shift);
              const unsigned long long mine = (unsigned long long)j << 24;
              const unsigned long long cur = hslot[s];
              if (cur != kEmpty && (cur >> 24) == (unsigned long long)j) {
                atomicAdd(hslot + s, 1ull);
                cached = true;
              } else if (cur == kEmpty) {
                const unsigned long long old = atomicCAS(hslot + s, kEmpty, mine | 1ull);
                if (old == kEmpty) {
                  cached = true;  // claimed with this deposit counted
                } else if ((old >> 24) == (unsigned long long)j) {
                  atomicAdd(hslot + s, 1ull);
                  cached = true;
                }
              }
            }
            if (RUNS3) {
              // the open run grows while the lane stays in its row at the next time
              const int jr = j < 0 ? -1 : (int)((uint32_t)sp - tbase + (isP ? 0u : p.tsp));
              if (cached) {
                if (jrow >= 0) flush_run();
              } else if (jrow >= 0 && jr == jrow && it == trun + lrun) {
                lrun += 1;
                if (lrun == 3) flush_run();
              } else {
                if (jrow >= 0) flush_run();
                if (jr >= 0) {
                  jrow = jr;
                  trun = it;
                  lrun = 1;
                  rsub = psub;
                }
              }
            } else if (cached) {
              if (tpend >= 0) __hip_atomic_fetch_add(PT + tpend, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              tpend = -1;
            } else if (tpend >= 0 && j == tpend + 1) {
              // the same counter row at two consecutive times: one 64-bit atomic on the
              // aligned pair, in T from an even time, in T2 (shifted by one) from an odd one
              RTMC_GLOBAL unsigned long long* w =
                  (tpend & 1) ? (RTMC_GLOBAL unsigned long long*)gl(P.T2) + ((tpend - 1) >> 1)
                              : (RTMC_GLOBAL unsigned long long*)PT + (tpend >> 1);
              __hip_atomic_fetch_add(w, 0x100000001ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              tpend = -1;
            } else {
              if (tpend >= 0) __hip_atomic_fetch_add(PT + tpend, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              tpend = -1;
              if (j >= 0) {
                if (it + 1 < p.n_times && (!(it & 1) || p.odd_pairs)) tpend = j;
                else __hip_atomic_fetch_add(PT + j, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              }
            }
            if (j >= 0) ndep += 1;
          } else if (idx >= 0) {
            ndep += 1;
            if (!ADJ) {
              if (LDSH) atomicAdd(shist + idx, 1u);
              else __hip_atomic_fetch_add(PE + idx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
              // score of the trajectory up to this observation: A = d log L / d log lam
              // (scattering tests), k = removal tests survived; summed through aw and
              // awi (fold)
              const double w = LDSH ? (double)sw[idx] : (double)P.w[idx];
              aw += w;
              awi += w * (double)it;
              if (P.dhist) {
                double A, k;
                if (QUANT) {
                  k = (double)it * (double)P.nsub;
                  A = (double)nsc * P.lam_over_expm1 - (k - (double)nsc) * P.lamS[0];
                } else {
                  k = 0.0;
                  A = (double)nsc - (double)it * p.dt_obs / M.mftP;
                }
                unsafeAtomicAdd(P.dhist + idx, A * P.two_over_c);
                unsafeAtomicAdd(P.dhist + P.dhist_stride + idx, A * P.dlogS + gsum);
                if (P.kq != 0.0) unsafeAtomicAdd(P.dhist + 2 * P.dhist_stride + idx, k * P.kq);
              }
            }
          }
          it += 1;
          if (it >= p.n_times) {
            has = false;
            need = true;
          }
        }
        }
      }
    }

    if (TMAJ && tpend >= 0) __hip_atomic_fetch_add(PT + tpend, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (RUNS3 && jrow >= 0) flush_run();
    fold();
    __syncthreads();
    if (TMAJ && !ADJ) {
      for (int k = threadIdx.x; k < p.hot_slots; k += blockDim.x) {
        const unsigned long long v = hslot[k];
        const uint32_t c = (uint32_t)(v & 0xFFFFFFull);
        if (v == kEmpty || !c) continue;
        if (RUNS3) {
          // straight into E (a slot counts up to 2^24 - 1, more than a 21-bit counter
          // holds): T index row * nt_pad + t (< 2^32, the host keeps these scratches
          // small), row = spatial - tbase + mode * tsp
          const uint32_t j = (uint32_t)(v >> 24);
          const uint32_t row = j / (uint32_t)p.nt_pad, t = j - row * (uint32_t)p.nt_pad;
          const uint32_t mode = row >= p.tsp ? 1u : 0u;
          const uint32_t sp = row - mode * p.tsp + (p.tcomp ? (uint32_t)bs.tdir * p.stride[3] : 0u);
          __hip_atomic_fetch_add(PE + ((int64_t)sp + (int64_t)t * p.time_stride + (mode ? p.mode_stride : 0)), c,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          __hip_atomic_fetch_add(PT + (v >> 24), c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    if (LDSH && !ADJ) {
      for (int64_t k = threadIdx.x; k < p.hist_size; k += blockDim.x) {
        uint32_t c = shist[k];
        if (c) __hip_atomic_fetch_add(PE + k, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (P.counters) {
      // one atomic per wavefront and counter (64-bit sums of the lanes' 32-bit counts)
      unsigned long long cs = nscat, cd = ndep, cb = nbnd, ck = nkill;
      for (int off = 32; off > 0; off >>= 1) {
        cs += __shfl_down(cs, off, 64);
        cd += __shfl_down(cd, off, 64);
        cb += __shfl_down(cb, off, 64);
        ck += __shfl_down(ck, off, 64);
      }
      if (lane == 0) {
        if (cs) atomicAdd(P.counters + RTMC_CNT_SCATTER, cs);
        if (cd) atomicAdd(P.counters + RTMC_CNT_DEPOSIT, cd);
        if (cb) atomicAdd(P.counters + RTMC_CNT_BOUNDARY, cb);
        if (ck) atomicAdd(P.counters + RTMC_CNT_KILLED, ck);
      }
    }
    if (ADJ) {
      for (int off = 32; off > 0; off >>= 1) {
        gaA += __shfl_down(gaA, off, 64);
        gaG += __shfl_down(gaG, off, 64);
        gaK += __shfl_down(gaK, off, 64);
      }
      if (lane == 0) {
        unsafeAtomicAdd(P.grad + 0, gaA * P.two_over_c);
        unsafeAtomicAdd(P.grad + 1, gaA * P.dlogS + gaG);
        unsafeAtomicAdd(P.grad + 2, gaK * P.kq);
      }
    }
    if (threadIdx.x == 0) atomicSub(p.active + cur, 1u);
    __syncthreads();  // the LDS tables and s_cur are rewritten for the next problem
  }
}

}  // namespace rtmc

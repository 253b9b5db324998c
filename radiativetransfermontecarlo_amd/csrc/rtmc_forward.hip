// rtmc_forward.hip -- the fused persistent kernel of rtmc_forward.
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
// scattering (or removal) is geometric; this kernel samples those indices
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
// Histogram: LDS-privatised per work-group when the whole histogram fits in
// LDS (flushed once per work-group with global atomics), otherwise global
// no-return atomics.
#include "rtmc_common.hpp"

#include <algorithm>
#include <cstdlib>

namespace rtmc {

constexpr int64_t kInf = (int64_t)1 << 60;

struct FwdParams {
  MatDev M;
  rtmc_geometry G;
  rtmc_source S;
  const double* edges[4];
  int ne[4];
  int eoff[4];        // offset of each edge array in the LDS copy
  int64_t stride[4];
  unsigned present;   // variables that exist (bit v; z is absent in 2-D)
  int64_t time_stride;
  int64_t mode_stride;
  int64_t n;          // particles of the launch
  uint64_t pid0;      // global id of particle 0
  uint32_t k0, k1;    // Philox key
  int n_times;
  int64_t nsub;       // sub-steps per observation interval
  double dt_sub;
  double dt_obs;
  double lamS[2];     // scattering: dt_s / mft(mode) (per sub-step)
  double lamK[2];     // removal: omega dt_s / Q(mode) (per sub-step)
  int useQ;
  int edges_total;    // doubles of the LDS edge copy (0: edges read from global memory)
  unsigned long long* work;  // global particle counter of the launch (zeroed)
  int chunk;          // particles fetched per counter access
  int fp32;           // flight arithmetic in FP32 (RTMC_PREC_FP32_FLIGHT)
  int batch;          // lanes waiting for an event before a wavefront takes an event step
  int ntab;           // inverse-CDF tables (1 acoustic, 4 elastic)
  int guide_n;        // guide cells per table (power of two; 0 = plain binary search)
  int64_t hist_size;  // uint32 words of the LDS histogram (0 = global atomics)
};

// Per-particle Philox stream: counter = {c, 0, pid_lo, pid_hi}, key = (k0, k1).
struct Rng {
  uint32_t k0, k1, plo, phi;
  uint32_t c;
  bool has;
  double spare;
  __device__ __forceinline__ double next() {
    if (has) {
      has = false;
      return spare;
    }
    U4 in;
    in.x = c;
    in.y = 0u;
    in.z = plo;
    in.w = phi;
    U4 r = philox10(in, k0, k1);
    c += 1u;
    spare = u53(r.z, r.w);
    has = true;
    return u53(r.x, r.y);
  }
  // uniform in (0, 1] with 32 random bits (FP32 flight): one Philox word each
  __device__ __forceinline__ double next32() {
    if (n32 == 0) {
      U4 in;
      in.x = c;
      in.y = 0u;
      in.z = plo;
      in.w = phi;
      w32 = philox10(in, k0, k1);
      c += 1u;
      n32 = 4;
    }
    uint32_t w = (n32 == 4) ? w32.x : (n32 == 3) ? w32.y : (n32 == 2) ? w32.z : w32.w;
    n32 -= 1;
    return ((double)w + 1.0) * (1.0 / 4294967296.0);
  }
  U4 w32;
  int n32;
  template <typename R>
  __device__ __forceinline__ double u() {
    if (sizeof(R) == 8) return next();
    return next32();
  }
};

// index of the next sub-step with a success, trials of probability 1-exp(-lam)
template <typename R>
__device__ __forceinline__ int64_t geom_after(Rng& r, double lam) {
  double u = r.u<R>();
  if (!(lam > 0.0)) return kInf;
  double e = (sizeof(R) == 8) ? -log(u) : (double)(-logf((float)u));
  double g = ceil(e / lam);
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
    double s0 = xb[b] - B.val;
    double r = h * d[b];
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
  double invw[4];
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

__device__ __forceinline__ void bin_setup(const FwdParams& p, int dim, const double* sedges, BinSetup* bs) {
  bs->check = 0u;
  for (int v = 0; v < 4; ++v) {
    bs->e0[v] = 0.0;
    bs->invw[v] = 0.0;
    if (!(p.present & (1u << v))) continue;
    const double* e = sedges ? sedges + p.eoff[v] : p.edges[v];
    double lo = e[0], hi = e[p.ne[v] - 1];
    if (p.ne[v] == 2 && trivially_inside(v, dim, p.G.frame, lo, hi)) continue;
    bs->check |= 1u << v;
    bs->e0[v] = lo;
    double w = hi - lo;
    bs->invw[v] = (isfinite(w) && w > 0.0) ? (double)(p.ne[v] - 1) / w : 0.0;
  }
}

template <int DIM>
__device__ __forceinline__ void birth_fwd64(const FwdParams& p, Rng& rng, double x[3], double d[3], double q[3]);

template <int DIM, typename R>
__device__ __forceinline__ void birth_fwd(const FwdParams& p, Rng& rng, R xo[3], R dd[3], R qo[3]) {
  double x[3], d[3], q[3];
  birth_fwd64<DIM>(p, rng, x, d, q);
  for (int k = 0; k < 3; ++k) {
    xo[k] = (R)x[k];
    dd[k] = (R)d[k];
    qo[k] = (R)q[k];
  }
}

// birth computed in FP64 for both flight precisions (cheap: once per particle;
// keeps dir_z = cos(pi/2) of 2-D births at its FP64 value)
template <int DIM>
__device__ __forceinline__ void birth_fwd64(const FwdParams& p, Rng& rng, double x[3], double d[3], double q[3]) {
  const rtmc_source& S = p.S;
  if (S.type == RTMC_SRC_POINT) {
    double r;
    if (S.radial.n == 0) {
      double u1 = rng.next(), u2 = rng.next();
      double g = sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
      r = fabs(g * S.lambda);  // initializeParticle.m:23
    } else {
      r = table_eval(S.radial.x, S.radial.y, S.radial.n, S.radial.extrap, rng.next());  // :25-27
    }
    double phi0, phi;
    if (DIM == 2) {
      phi0 = kPi / 2;
      phi = kPi / 2;
    } else {
      phi0 = acos(-1.0 + 2.0 * rng.next());
      phi = acos(-1.0 + 2.0 * rng.next());
    }
    double theta0 = kTwoPi * rng.next();
    double theta = kTwoPi * rng.next();
    if (S.direction == RTMC_DIR_OUTGOING) {
      theta = theta0;
      phi = phi0;
    } else if (S.direction == RTMC_DIR_UPPER) {
      theta = kTwoPi * rng.next();
      phi = acos(rng.next());
    }
    double st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi, &sp, &cp);
    d[0] = ct * sp;
    d[1] = st * sp;
    d[2] = cp;
    q[0] = -st;
    q[1] = ct;
    q[2] = 0.0;
    if (DIM == 3) {
      double alpha = kTwoPi * rng.next();
      double c[3];
      cross3(d, q, c);
      double sa, ca;
      sincos(alpha, &sa, &ca);
      for (int k = 0; k < 3; ++k) q[k] = q[k] * ca + c[k] * sa;
    }
    double st0, ct0, sp0, cp0;
    sincos(theta0, &st0, &ct0);
    sincos(phi0, &sp0, &cp0);
    x[0] = S.position[0] + r * (ct0 * sp0);
    x[1] = S.position[1] + r * (st0 * sp0);
    x[2] = S.position[2] + r * cp0;
  } else {
    double u1 = rng.next(), u2 = rng.next();
    double g = sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
    double c_long = g * S.lambda / 2.0;
    double c1, c2;
    if (S.aperture == RTMC_APERTURE_RECTANGLE) {
      c1 = (rng.next() - 0.5) * S.extent[0];
      c2 = (rng.next() - 0.5) * S.extent[1];
    } else {
      double rho = S.extent[0] * sqrt(rng.next());
      double alpha = kTwoPi * rng.next();
      double sa, ca;
      sincos(alpha, &sa, &ca);
      c1 = rho * ca;
      c2 = rho * sa;
    }
    for (int k = 0; k < 3; ++k)
      x[k] = ((S.position[k] + c_long * S.normal[k]) + c1 * S.basis_u[k]) + c2 * S.basis_v[k];
    for (int k = 0; k < 3; ++k) d[k] = S.normal[k];
    double th = rng.next() * kTwoPi;
    double s, c;
    sincos(th, &s, &c);
    for (int k = 0; k < 3; ++k) q[k] = c * S.basis_u[k] + s * S.basis_v[k];
  }
}

template <int DIM, bool LDSH, typename R>
__device__ __forceinline__ void deposit(const FwdParams& p, const BinSetup& bs, const double* __restrict__ sedges,
                                        uint32_t* __restrict__ shist, uint32_t* __restrict__ E, const R X[3], R dz,
                                        int isP, int it, unsigned long long& ndep) {
  double v[4];
  obs_vars(DIM, p.G.frame, X, dz, bs.check, v);
  int64_t idx = (int64_t)it * p.time_stride;
  if (p.edges_total > 0) {
    // edges in LDS (`sedges` points into the work-group's LDS block)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!(bs.check & (1u << k))) continue;
      int b = bin_guess(sedges + p.eoff[k], p.ne[k], bs.e0[k], bs.invw[k], v[k]);
      if (b < 0) return;
      idx += (int64_t)b * p.stride[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!(bs.check & (1u << k))) continue;
      int b = bin_guess(p.edges[k], p.ne[k], bs.e0[k], bs.invw[k], v[k]);
      if (b < 0) return;
      idx += (int64_t)b * p.stride[k];
    }
  }
  if (!isP) idx += p.mode_stride;
  ndep += 1;
  if (LDSH) {
    atomicAdd(shist + idx, 1u);
  } else {
    __hip_atomic_fetch_add(E + idx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// LDS-histogram kernels run one 1024-lane work-group per CU (the histogram
// takes most of the LDS); the others use 256-lane work-groups.
template <int DIM, bool LDSH, bool QUANT, bool AC, bool BND, typename R>
__global__ __launch_bounds__(LDSH ? 1024 : 256) void k_forward(const FwdParams p, uint32_t* __restrict__ E,
                                                 unsigned long long* __restrict__ counters) {
  extern __shared__ double smem[];
  double* sedges = smem;
  __shared__ BinSetup s_bs;
  __shared__ double s_tx0[4], s_tx1[4];
  int* sguide = (int*)(smem + p.edges_total);
  uint32_t* shist = (uint32_t*)(sguide + (p.guide_n > 0 ? p.ntab * (p.guide_n + 1) : 0));
  if (p.guide_n > 0) {
    // guide table: cell c holds the largest knot index k <= n-2 with x[k] <= c/G
    // (Chen & Asau's guide for inversion of a tabulated CDF)
    const int G = p.guide_n;
    for (int t = 0; t < p.ntab; ++t) {
      const TableDev& T = p.M.inv[t];
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
  if (p.edges_total > 0) {
    // gather the edge arrays into one LDS block
    for (int v = 0; v < 4; ++v) {
      if (!(p.present & (1u << v))) continue;
      for (int k = threadIdx.x; k < p.ne[v]; k += blockDim.x) smem[p.eoff[v] + k] = p.edges[v][k];
    }
  }
  if (LDSH) {
    for (int64_t k = threadIdx.x; k < p.hist_size; k += blockDim.x) shist[k] = 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) bin_setup(p, DIM, sedges, &s_bs);
  __syncthreads();
  const BinSetup bs = s_bs;

  const MatDev& M = p.M;
  // inverse-CDF lookup of the reference table (same interval and formula as
  // table_eval), bracketed by the LDS guide when present
  auto lookup = [&](int t, double u) -> double {
    const TableDev& T = M.inv[t];
    if (p.guide_n > 0 && u >= s_tx0[t] && u <= s_tx1[t]) {
      const int G = p.guide_n;
      int c = (int)(u * (double)G);
      c = c < 0 ? 0 : (c > G - 1 ? G - 1 : c);
      const int* g = sguide + t * (G + 1);
      return table_eval_bracket(T.x, T.y, T.n, T.extrap, u, g[c], (int64_t)g[c + 1] + 1);
    }
    return table_eval(T.x, T.y, T.n, T.extrap, u);
  };
  unsigned long long nscat = 0, ndep = 0, nbnd = 0, nkill = 0;
  // ---- lane state machine ------------------------------------------------
  // Each lane holds one particle and its own next observation index.  One
  // iteration is either a DEPOSIT step (every flying lane deposits its next
  // observation) or an EVENT step (lanes whose next event comes before that
  // observation process it, idle lanes take a new particle).  The wavefront
  // takes an event step only when at least `batch` lanes wait for one (or no
  // lane can fly), so the expensive event path runs for many lanes at once
  // instead of once per observation for whichever lane happens to need it.
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  Rng rng;
  rng.k0 = p.k0;
  rng.k1 = p.k1;
  rng.plo = 0u;
  rng.phi = 0u;
  rng.c = 0u;
  rng.has = false;
  rng.spare = 0.0;
  rng.n32 = 0;
  R xb[3] = {R(0), R(0), R(0)}, d[3] = {R(0), R(0), R(1)}, q[3] = {R(1), R(0), R(0)};
  int isP = 1;
  R h = R(0);                     // sub-stepped: step length v dt_s; event-driven: speed v
  int64_t kb = 0, Ks = kInf, Kk = kInf, kev = kInf;   // sub-stepped: base, next scatter / removal / event
  double tb = 0.0, ts = INFINITY;                     // event-driven: base time, next scattering time
  int it = 1;                     // next observation index
  bool has = false;               // lane holds a particle
  bool need = true;               // lane wants a particle
  // wave-uniform pool of particle indices fetched from the global counter
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
    const bool pend_ev = has && (QUANT ? (kev <= (int64_t)it * p.nsub) : (ts <= (double)it * p.dt_obs));
    const bool pend = need || pend_ev;
    const bool fly = has && !pend_ev;
    const uint64_t pm = __ballot(pend);
    const uint64_t fm = __ballot(fly);
    if (pm == 0ull && fm == 0ull) break;
    if (__popcll(pm) >= p.batch || fm == 0ull) {
      // ---------------- EVENT step ----------------
      const uint64_t nm = __ballot(need);
      if (nm) {
        // hand out particle indices to the lanes that need one
        const int n = __popcll(nm);
        int64_t L = pool_end - pool_next;
        int64_t nb0 = 0, nb1 = 0;
        if (L < n && !exhausted) {
          unsigned long long base = 0;
          if (lane == __ffsll((long long)nm) - 1) base = atomicAdd(p.work, (unsigned long long)p.chunk);
          base = __shfl(base, __ffsll((long long)nm) - 1, 64);
          if ((int64_t)base >= p.n) {
            exhausted = true;
          } else {
            nb0 = (int64_t)base;
            nb1 = min((int64_t)base + (int64_t)p.chunk, p.n);
          }
        }
        if (need) {
          const int r = __popcll(nm & lt_mask);
          int64_t id = -1;
          if (r < L) id = pool_next + r;
          else if (r - L < nb1 - nb0) id = nb0 + (r - L);
          if (id >= 0) {
            const uint64_t pid = p.pid0 + (uint64_t)id;
            rng.plo = (uint32_t)pid;
            rng.phi = (uint32_t)(pid >> 32);
            rng.c = 0u;
            rng.has = false;
            rng.n32 = 0;
            birth_fwd<DIM>(p, rng, xb, d, q);
            isP = (!AC && p.S.polarization == 1) ? 0 : 1;
            it = 1;
            has = true;
            need = false;
            if (QUANT) {
              h = (R)((isP ? M.vp : M.vs) * p.dt_sub);
              kb = 0;
              Ks = geom_after<R>(rng, p.lamS[isP ? 0 : 1]);
              Kk = p.useQ ? geom_after<R>(rng, p.lamK[isP ? 0 : 1]) : kInf;
              next_event();
            } else {
              h = (R)(isP ? M.vp : M.vs);
              tb = 0.0;
              ts = -log(rng.u<R>()) * (isP ? M.mftP : M.mftS);
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
              if (!AC) {
                uA = rng.u<R>();
                uB = rng.u<R>();
              }
              if (B.dir < 4) plane_reflect(M, AC, B, x, d, q, isP, uA, uB);
              else cyl_reflect(M, AC, B, x, d, q, isP, uA, uB);
            }
            if (alive) {
              if (!AC && isP != wasP) {
                // mode changed at the boundary: this sub-step's scattering test and
                // the removal tests from the next one use the new mode
                h = (R)((isP ? M.vp : M.vs) * p.dt_sub);
                Ks = j - 1 + geom_after<R>(rng, p.lamS[isP ? 0 : 1]);
                Kk = p.useQ ? j + geom_after<R>(rng, p.lamK[isP ? 0 : 1]) : kInf;
              }
              if (Ks == j) {
                const int before = isP;
                double uphi = rng.u<R>();
                double uth = rng.u<R>();
                double uconv = AC ? 0.0 : rng.u<R>();
                double th = scatter_angles_t<DIM>(M, AC, d, q, isP, uphi < 0.5, uphi, uth, uconv, lookup);
                rotate_renorm(d, q, (R)th);
                nscat += 1;
                if (!AC && isP != before) {
                  h = (R)((isP ? M.vp : M.vs) * p.dt_sub);
                  if (p.useQ) Kk = j + geom_after<R>(rng, p.lamK[isP ? 0 : 1]);
                }
                Ks = j + geom_after<R>(rng, p.lamS[isP ? 0 : 1]);
              }
              xb[0] = x[0];
              xb[1] = x[1];
              xb[2] = x[2];
              kb = j;
              next_event();
            }
          }
          if (!alive) {
            has = false;
            need = true;
          }
        } else {
          // event-driven: scattering at time ts (propagateParticle.m:24-72)
          const R f = (R)((double)h * (ts - tb));
          for (int k = 0; k < 3; ++k) xb[k] = xb[k] + f * d[k];
          tb = ts;
          double uphi = rng.u<R>();
          double uth = rng.u<R>();
          double uconv = AC ? 0.0 : rng.u<R>();
          double th = scatter_angles_t<DIM>(M, AC, d, q, isP, uphi < 0.5, uphi, uth, uconv, lookup);
          rotate_renorm(d, q, (R)th);
          nscat += 1;
          h = (R)(isP ? M.vp : M.vs);
          ts = tb - log(rng.u<R>()) * (isP ? M.mftP : M.mftS);
        }
      }
    } else {
      // ---------------- DEPOSIT step ----------------
      if (fly) {
        R X[3];
        if (QUANT) {
          const R m = (R)((int64_t)it * p.nsub - kb);
          for (int k = 0; k < 3; ++k) X[k] = xb[k] + (m * h) * d[k];
        } else {
          const R f = (R)((double)h * ((double)it * p.dt_obs - tb));
          for (int k = 0; k < 3; ++k) X[k] = xb[k] + f * d[k];
        }
        deposit<DIM, LDSH>(p, bs, sedges, shist, E, X, d[2], isP, it, ndep);
        it += 1;
        if (it >= p.n_times) {
          has = false;
          need = true;
        }
      }
    }
  }

  if (LDSH) {
    __syncthreads();
    for (int64_t k = threadIdx.x; k < p.hist_size; k += blockDim.x) {
      uint32_t c = shist[k];
      if (c) __hip_atomic_fetch_add(E + k, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (counters) {
    // one atomic per wavefront and counter
    for (int off = 32; off > 0; off >>= 1) {
      nscat += __shfl_down(nscat, off, 64);
      ndep += __shfl_down(ndep, off, 64);
      nbnd += __shfl_down(nbnd, off, 64);
      nkill += __shfl_down(nkill, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      if (nscat) atomicAdd(counters + RTMC_CNT_SCATTER, nscat);
      if (ndep) atomicAdd(counters + RTMC_CNT_DEPOSIT, ndep);
      if (nbnd) atomicAdd(counters + RTMC_CNT_BOUNDARY, nbnd);
      if (nkill) atomicAdd(counters + RTMC_CNT_KILLED, nkill);
    }
  }
}

// ------------------------------------------------------------------
// host side

template <int DIM, bool LDSH, bool QUANT, bool AC, bool BND, typename R>
static int launch_fwd(const FwdParams& p, size_t smem, unsigned grid, unsigned block, uint32_t* E,
                      unsigned long long* counters, hipStream_t st) {
  auto kern = k_forward<DIM, LDSH, QUANT, AC, BND, R>;
  if (smem > 64 * 1024) RTMC_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)smem));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), smem, st, p, E, counters);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

// compile-time specialisation: dimension, histogram placement, process,
// acoustic/elastic, with/without boundaries (the event-driven process has none)
template <int DIM, bool LDSH, typename R>
static int dispatch_fwd(const FwdParams& p, bool quant, bool ac, bool bnd, size_t smem, unsigned grid,
                        unsigned block, uint32_t* E, unsigned long long* counters, hipStream_t st) {
  if (quant) {
    if (ac) {
      if (bnd) return launch_fwd<DIM, LDSH, true, true, true, R>(p, smem, grid, block, E, counters, st);
      return launch_fwd<DIM, LDSH, true, true, false, R>(p, smem, grid, block, E, counters, st);
    }
    if (bnd) return launch_fwd<DIM, LDSH, true, false, true, R>(p, smem, grid, block, E, counters, st);
    return launch_fwd<DIM, LDSH, true, false, false, R>(p, smem, grid, block, E, counters, st);
  }
  if (ac) return launch_fwd<DIM, LDSH, false, true, false, R>(p, smem, grid, block, E, counters, st);
  return launch_fwd<DIM, LDSH, false, false, false, R>(p, smem, grid, block, E, counters, st);
}

template <typename R>
static int dispatch_prec(const FwdParams& p, int dim, bool ldsh, bool quant, bool ac, bool bnd, size_t smem,
                         unsigned grid, unsigned block, uint32_t* E, unsigned long long* cnt, hipStream_t st) {
  if (dim == 2) {
    if (ldsh) return dispatch_fwd<2, true, R>(p, quant, ac, bnd, smem, grid, block, E, cnt, st);
    return dispatch_fwd<2, false, R>(p, quant, ac, bnd, smem, grid, block, E, cnt, st);
  }
  if (ldsh) return dispatch_fwd<3, true, R>(p, quant, ac, bnd, smem, grid, block, E, cnt, st);
  return dispatch_fwd<3, false, R>(p, quant, ac, bnd, smem, grid, block, E, cnt, st);
}

}  // namespace rtmc

using namespace rtmc;

static int launch_all(const FwdParams& p, int dim, bool ldsh, bool quant, bool ac, bool bnd, size_t smem,
                      unsigned grid, unsigned block, uint32_t* E, unsigned long long* cnt, hipStream_t st);

extern "C" int rtmc_forward(const rtmc_problem* prob, uint32_t* E, uint64_t* counters, void* stream) {
  clear_error();
  RTMC_REQUIRE(prob != nullptr && E != nullptr, "problem and E must be given");
  int e = check_geometry(&prob->geometry);
  if (e) return e;
  e = check_material(&prob->material, &prob->geometry);
  if (e) return e;
  const int dim = prob->geometry.dimension;
  RTMC_REQUIRE(prob->n_particles >= 0, "n_particles must be >= 0");
  RTMC_REQUIRE(prob->n_times >= 1, "n_times must be >= 1");
  RTMC_REQUIRE(prob->time_steps == 0 || prob->time_steps == 1, "material.timeSteps must be 0 or 1");
  RTMC_REQUIRE(prob->precision == RTMC_PREC_FP64 || prob->precision == RTMC_PREC_FP32_FLIGHT,
               "precision must be RTMC_PREC_FP64 or RTMC_PREC_FP32_FLIGHT");
  RTMC_REQUIRE(prob->source.type == RTMC_SRC_POINT || dim == 3, "plane sources are 3-D only");
  if (prob->source.type == RTMC_SRC_POINT && prob->source.radial.n != 0) {
    e = check_table(&prob->source.radial, "source.radial");
    if (e) return e;
  }
  if (prob->time_steps == 1) {
    // propagateParticle has neither Q nor boundaries (propagateParticle.m:1-81)
    RTMC_REQUIRE(prob->geometry.n_bnd == 0, "timeSteps = 1 (propagateParticle) does not handle boundaries");
  }
  if (prob->n_particles == 0 || prob->n_times < 2) return RTMC_OK;
  hipStream_t st = (hipStream_t)stream;

  FwdParams p;
  p.M = to_matdev(&prob->material);
  p.G = prob->geometry;
  p.S = prob->source;
  const rtmc_bins& B = prob->bins;
  int total = 0;
  p.present = 0;
  for (int v = 0; v < 4; ++v) {
    p.edges[v] = B.edges[v];
    p.ne[v] = B.n_edges[v];
    p.stride[v] = B.stride[v];
    p.eoff[v] = total;
    if (dim == 2 && v == 2) continue;
    RTMC_REQUIRE(B.n_edges[v] >= 2 && B.edges[v] != nullptr, "variable %d needs >= 2 edges", v + 1);
    p.present |= 1u << v;
    total += B.n_edges[v];
  }
  p.time_stride = B.time_stride;
  p.mode_stride = B.mode_stride;
  p.n = prob->n_particles;
  p.pid0 = prob->particle_offset;
  uint64_t key = prob->seed ^ 0x52544D4346574400ull;  // forward-kernel stream family
  p.k0 = (uint32_t)key;
  p.k1 = (uint32_t)(key >> 32);
  p.n_times = prob->n_times;
  p.dt_obs = prob->dt;

  // sub-step of propagateParticleSmallDt.m:4-7 with T - mean(P.t) = dt
  const rtmc_material& m = prob->material;
  double mftP = m.mft[0], mftS = m.acoustics ? m.mft[0] : m.mft[1];
  double mftmin = m.acoustics ? mftP : fmin(mftP, mftS);
  double Te = prob->dt;
  double dts = fmin(0.01 * mftmin, Te);
  int64_t nsub = (Te > 0.0) ? (int64_t)ceil(Te / dts) : 1;
  RTMC_REQUIRE(nsub >= 1 && nsub < ((int64_t)1 << 40), "sub-step count out of range");
  dts = Te / (double)nsub;
  p.nsub = nsub;
  p.dt_sub = dts;
  p.lamS[0] = dts / mftP;
  p.lamS[1] = dts / mftS;
  double QP = m.QP, QS = m.acoustics ? m.QP : m.QS;
  p.useQ = (prob->time_steps == 0) && (std::isfinite(QP) || std::isfinite(QS));
  double omega = 2.0 * kPi * m.frequency;
  p.lamK[0] = (p.useQ && std::isfinite(QP)) ? omega * dts / QP : 0.0;
  p.lamK[1] = (p.useQ && std::isfinite(QS)) ? omega * dts / QS : 0.0;

  // histogram extent: all time slices and modes
  int64_t hsize = B.time_stride * prob->n_times;
  if (!m.acoustics) hsize = std::max(hsize, 2 * B.mode_stride);

  int dev = 0;
  RTMC_HIP(hipGetDevice(&dev));
  int ncu = 0, max_smem = 0;
  RTMC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  RTMC_HIP(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));

  size_t edge_bytes = (size_t)total * sizeof(double);
  bool edges_lds = edge_bytes <= 32 * 1024;
  p.edges_total = edges_lds ? total : 0;
  if (!edges_lds) edge_bytes = 0;
  size_t hist_bytes = (size_t)hsize * sizeof(uint32_t);
  p.ntab = m.acoustics ? 1 : 4;
  int guide = 2048;
  if (const char* g = getenv("RTMC_GUIDE_CELLS")) guide = atoi(g);
  RTMC_REQUIRE(guide >= 0 && guide <= (1 << 16) && (guide & (guide - 1)) == 0,
               "RTMC_GUIDE_CELLS must be 0 or a power of two <= 65536");
  auto guide_bytes = [&](int G) { return G > 0 ? (size_t)p.ntab * (size_t)(G + 1) * sizeof(int) : (size_t)0; };
  bool ldsh;
  if (prob->hist_mode == 1) ldsh = false;
  else if (prob->hist_mode == 2) ldsh = true;
  else ldsh = (edge_bytes + hist_bytes) <= (size_t)max_smem && hsize <= (int64_t)1 << 20;
  if (ldsh && edge_bytes + hist_bytes > (size_t)max_smem) {
    set_error("LDS histogram of %zu bytes exceeds the %d bytes of LDS per work-group", hist_bytes, max_smem);
    return RTMC_ERR_ARG;
  }
  // the guide takes what the histogram leaves
  while (guide > 0 && edge_bytes + guide_bytes(guide) + (ldsh ? hist_bytes : 0) > (size_t)max_smem) guide >>= 1;
  if (guide < 64) guide = 0;
  p.guide_n = guide;
  p.hist_size = ldsh ? hsize : 0;
  size_t smem = edge_bytes + guide_bytes(guide) + (ldsh ? hist_bytes : 0);

  unsigned block = prob->block_size > 0 ? (unsigned)prob->block_size : (ldsh ? 1024u : 256u);
  RTMC_REQUIRE(block % 64 == 0 && block <= (ldsh ? 1024u : 256u),
               "block_size must be a multiple of 64 and <= 256 (<= 1024 with the LDS histogram)");
  int64_t want = (prob->n_particles + block - 1) / block;
  int64_t per_cu;
  if (ldsh) per_cu = std::max<int64_t>(1, (int64_t)max_smem / (int64_t)std::max<size_t>(smem, 1)) ;
  else per_cu = 2048 / block;
  if (prob->particles_per_thread > 0)
    want = (prob->n_particles + (int64_t)block * prob->particles_per_thread - 1) /
           ((int64_t)block * prob->particles_per_thread);
  int64_t cap = (int64_t)ncu * per_cu;
  unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, cap));

  // dynamic work distribution: one zeroed counter per launch
  p.chunk = 256;
  p.fp32 = prob->precision == RTMC_PREC_FP32_FLIGHT;
  p.batch = 24;
  if (const char* b = getenv("RTMC_EVENT_BATCH")) p.batch = atoi(b);
  RTMC_REQUIRE(p.batch >= 1 && p.batch <= 64, "RTMC_EVENT_BATCH must be in [1, 64]");
  void* work = nullptr;
  RTMC_HIP(hipMallocAsync(&work, sizeof(unsigned long long), st));
  RTMC_HIP(hipMemsetAsync(work, 0, sizeof(unsigned long long), st));
  p.work = (unsigned long long*)work;
  int rc = launch_all(p, dim, ldsh, prob->time_steps == 0, m.acoustics != 0, prob->geometry.n_bnd > 0, smem, grid,
                      block, E, (unsigned long long*)counters, st);
  hipError_t fe = hipFreeAsync(work, st);
  if (rc == RTMC_OK && fe != hipSuccess) {
    set_error("hipFreeAsync failed: %s", hipGetErrorString(fe));
    return RTMC_ERR_HIP;
  }
  return rc;
}

static int launch_all(const FwdParams& p, int dim, bool ldsh, bool quant, bool ac, bool bnd, size_t smem,
                      unsigned grid, unsigned block, uint32_t* E, unsigned long long* cnt, hipStream_t st) {
  if (p.fp32) return dispatch_prec<float>(p, dim, ldsh, quant, ac, bnd, smem, grid, block, E, cnt, st);
  return dispatch_prec<double>(p, dim, ldsh, quant, ac, bnd, smem, grid, block, E, cnt, st);
}

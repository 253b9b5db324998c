// rtmc_mirror.hip -- function-level kernels of librtmc.
//
// One kernel per reference function, operating on a packet of particles
// held in device memory as structure-of-arrays (the column-major N x 3
// layout of the MATLAB struct P).  One thread owns one particle; the
// reference's vector operations over the packet become per-particle code,
// which is exact because particles never interact.  These kernels accept
// recorded variates (REPLAY) for bit-level parity runs, or Philox variates.
// The throughput path is the fused kernel of rtmc_forward.hip.
//
// Compiled without FMA contraction so that the arithmetic follows the
// reference's operation order.
#include "rtmc_common.hpp"

namespace rtmc {

constexpr int kBlock = 256;

__device__ __forceinline__ void ld3(const double* A, int64_t n, int64_t i, double v[3]) {
  v[0] = A[i];
  v[1] = A[n + i];
  v[2] = A[2 * n + i];
}
__device__ __forceinline__ void st3(double* A, int64_t n, int64_t i, const double v[3]) {
  A[i] = v[0];
  A[n + i] = v[1];
  A[2 * n + i] = v[2];
}

// ------------------------------------------------------------------
// initializeParticle (initializeParticle.m:1-171).  `U(slot)` returns the
// uniform of a draw site; `G()` the standard normal of slot 0.
template <typename UF, typename GF>
__device__ __forceinline__ void birth(const rtmc_source& S, int dim, UF U, GF G, double x[3], double dir[3],
                                      double perp[3]) {
  if (S.type == RTMC_SRC_POINT) {
    double r;
    if (S.radial.n == 0) {
      r = fabs(G() * S.lambda);  // :23
    } else {
      r = table_eval(S.radial.x, S.radial.y, S.radial.n, S.radial.extrap, U(0));  // :25-27
    }
    double phi0, phi;
    if (dim == 2) {  // :31-33
      phi0 = kPi / 2;
      phi = kPi / 2;
    } else {  // :35-36
      phi0 = acos(-1.0 + 2.0 * U(1));
      phi = acos(-1.0 + 2.0 * U(2));
    }
    double theta0 = 2.0 * kPi * U(3);  // :38
    double theta = 2.0 * kPi * U(4);   // :39
    if (S.direction == RTMC_DIR_OUTGOING) {  // :40-42
      theta = theta0;
      phi = phi0;
    } else if (S.direction == RTMC_DIR_UPPER) {  // :44-46
      theta = 2.0 * kPi * U(5);
      phi = acos(U(6));
    }
    dir[0] = cos(theta) * sin(phi);  // :48
    dir[1] = sin(theta) * sin(phi);
    dir[2] = cos(phi);
    perp[0] = -sin(theta);  // :49
    perp[1] = cos(theta);
    perp[2] = 0.0;
    if (dim == 3) {  // :51-54
      double alpha = 2.0 * kPi * U(7);
      double c[3];
      cross3(dir, perp, c);
      double ca = cos(alpha), sa = sin(alpha);
      for (int k = 0; k < 3; ++k) perp[k] = perp[k] * ca + c[k] * sa;
    }
    double ux = cos(theta0) * sin(phi0), uy = sin(theta0) * sin(phi0), uz = cos(phi0);  // :57-58
    x[0] = S.position[0] + r * ux;
    x[1] = S.position[1] + r * uy;
    x[2] = S.position[2] + r * uz;
  } else {
    // plane source (:61-147); n and null(n) prepared by the host (:67-100)
    double c_long = G() * S.lambda / 2.0;  // :94
    double c1, c2;
    if (S.aperture == RTMC_APERTURE_RECTANGLE) {  // :104-119
      c1 = (U(1) - 0.5) * S.extent[0];
      c2 = (U(2) - 0.5) * S.extent[1];
    } else {  // :121-136
      double rho = S.extent[0] * sqrt(U(1));
      double alpha = 2.0 * kPi * U(2);
      c1 = rho * cos(alpha);
      c2 = rho * sin(alpha);
    }
    for (int k = 0; k < 3; ++k)
      x[k] = ((S.position[k] + c_long * S.normal[k]) + c1 * S.basis_u[k]) + c2 * S.basis_v[k];
    for (int k = 0; k < 3; ++k) dir[k] = S.normal[k];  // :143
    double th = U(4) * 2.0 * kPi;                      // :146
    for (int k = 0; k < 3; ++k) perp[k] = cos(th) * S.basis_u[k] + sin(th) * S.basis_v[k];
  }
}

__global__ __launch_bounds__(kBlock) void k_init(const rtmc_source S, int acoustics, const SlotDraws D,
                                                rtmc_particles P) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  auto U = [&](int slot) { return D(i, slot, 0); };
  auto G = [&]() {
    if (D.kind == RTMC_DRAWS_REPLAY) return D(i, 0, 0);
    double u1 = D(i, 0, 0), u2 = D(i, 8, 0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
  };
  double x[3], dir[3], perp[3];
  birth(S, P.dimension, U, G, x, dir, perp);
  st3(P.x, n, i, x);
  st3(P.dir, n, i, dir);
  st3(P.perp, n, i, perp);
  P.t[i] = 0.0;                                                    // :152
  P.p[i] = (!acoustics && S.polarization == 1) ? 0 : 1;            // :157-160
  P.alive[i] = 1;                                                  // :170
}

// ------------------------------------------------------------------
// propagateParticleSmallDt, sub-step loop of one particle
// (propagateParticleSmallDt.m:85-642).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_smalldt(const MatDev M, const rtmc_geometry G, rtmc_particles P,
                                                   double dt, int64_t nsub, double pKP, double pKS, int useQ,
                                                   const SlotDraws D, double Te) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  const bool replay = D.kind == RTMC_DRAWS_REPLAY;
  int alive = P.alive[i];
  int isP = P.p[i];
  double x[3], d[3], q[3];
  ld3(P.x, n, i, x);
  ld3(P.dir, n, i, d);
  ld3(P.perp, n, i, q);
  for (int64_t i1 = 0; i1 < nsub && alive; ++i1) {
    if (useQ) {  // :100-138
      double pk = (M.acoustics || isP) ? pKP : pKS;
      if (pk > 0) {
        if (D(i, 0, i1) < pk) {
          alive = 0;
          break;
        }
      }
    }
    double x1[3] = {x[0], x[1], x[2]};  // :141-143
    double h = (isP ? M.vp : M.vs) * dt;
    for (int k = 0; k < 3; ++k) x[k] = x[k] + h * d[k];
    for (int ib = 0; ib < G.n_bnd && alive; ++ib) {  // :146-582
      const rtmc_boundary& B = G.bnd[ib];
      bool hit = (B.dir < 4) ? plane_crossed(B, x, x1) : cyl_outside(B, x);
      if (!hit) continue;
      if (B.type == RTMC_BND_ABSORBING) {
        alive = 0;
        break;
      }
      double uA = 0.0, uB = 0.0;
      if (!M.acoustics) {
        uA = D(i, 6 + 2 * ib, i1);
        uB = D(i, 7 + 2 * ib, i1);
      }
      if (B.dir < 4) plane_reflect(M, M.acoustics != 0, B, x, d, q, isP, uA, uB);
      else cyl_reflect(M, M.acoustics != 0, B, x, d, q, isP, uA, uB);
    }
    if (!alive) break;
    double lam = dt / (isP ? M.mftP : M.mftS);  // :591-601
    double us = D(i, 1, i1);
    bool scat = replay ? (us > 0.0) : (us < 1.0 - exp(-lam));
    if (scat) {  // :604-632
      double ua = D(i, 2, i1);
      bool bit = replay ? (ua != 0.0) : (ua < 0.5);
      double uth = D(i, 3, i1);
      double uconv = M.acoustics ? 0.0 : (isP ? D(i, 4, i1) : D(i, 5, i1));
      double th = scatter_angles<DIM>(M, d, q, isP, bit, ua, uth, uconv);
      rotate_renorm(d, q, th);  // :635-639
    } else {
      renorm(d, q);
    }
  }
  st3(P.x, n, i, x);
  st3(P.dir, n, i, d);
  st3(P.perp, n, i, q);
  P.p[i] = (uint8_t)isP;
  P.alive[i] = (uint8_t)alive;
  P.t[i] = P.t[i] + Te;  // :645
}

// Fixed-order sum of P.t (for mean(P.t), propagateParticleSmallDt.m:4).
__global__ __launch_bounds__(1024) void k_sum(const double* __restrict__ t, int64_t n, double* out) {
  __shared__ double s[1024];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) acc += t[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// ------------------------------------------------------------------
// propagateParticle (propagateParticle.m:1-96), loop of one particle.
__device__ __forceinline__ int64_t poisson_inv(double lam, double u) {
  int64_t k = 0;
  double pk = exp(-lam), F = pk;
  while (u > F && pk > 0.0) {
    k += 1;
    pk = pk * lam / (double)k;
    F += pk;
  }
  return k;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_events(const MatDev M, rtmc_particles P, double T, const SlotDraws D,
                                                  int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  const bool replay = D.kind == RTMC_DRAWS_REPLAY;
  int isP = P.p[i];
  double x[3], d[3], q[3];
  ld3(P.x, n, i, x);
  ld3(P.dir, n, i, d);
  ld3(P.perp, n, i, q);
  double t = P.t[i];
  double dt = T - t;  // :4
  int64_t it = 0;
  while (dt > 0) {  // :8
    if (replay && it >= D.n_steps) {
      atomicOr(flag, 1);
      break;
    }
    double mft = isP ? M.mftP : M.mftS;
    double u0 = D(i, 0, it);
    int64_t Nj;
    if (replay) {
      Nj = (int64_t)u0;
    } else {
      double lam = dt / mft;
      if (lam > 700.0) {
        atomicOr(flag, 2);
        break;
      }
      Nj = poisson_inv(lam, u0);
    }
    bool sc = Nj > 0;  // :20
    if (sc) {          // :24-26
      double umin = D(i, 1, it);
      if (!replay) umin = 1.0 - pow(umin, 1.0 / (double)Nj);
      dt = umin * dt;
    }
    double h = (isP ? M.vp : M.vs) * dt;  // :29-30
    for (int k = 0; k < 3; ++k) x[k] = x[k] + h * d[k];
    if (sc) {
      double ua = D(i, 2, it);
      bool bit = replay ? (ua != 0.0) : (ua < 0.5);
      double uth = D(i, 3, it);
      double uconv = M.acoustics ? 0.0 : (isP ? D(i, 4, it) : D(i, 5, it));
      double th = scatter_angles<DIM>(M, d, q, isP, bit, ua, uth, uconv);
      rotate_renorm(d, q, th);  // :68-72
    }
    t = t + dt;  // :75
    dt = T - t;  // :76
    it += 1;
  }
  st3(P.x, n, i, x);
  st3(P.dir, n, i, d);
  st3(P.perp, n, i, q);
  P.p[i] = (uint8_t)isP;
  P.t[i] = t;
}

// ------------------------------------------------------------------
// observeTime + E(:,:,it,:) += ... (observeTime.m:1-146, radiativeTransfer.m:75-76)
__global__ __launch_bounds__(kBlock) void k_observe(const rtmc_geometry G, const rtmc_bins B, int acoustics,
                                                   const rtmc_particles P, uint32_t* __restrict__ E,
                                                   int64_t time_index) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (!P.alive[i]) return;
  const int64_t n = P.n;
  double X[3];
  ld3(P.x, n, i, X);
  double v[4];
  double dz = P.dir[2 * n + i];
  if (B.direction_variable == RTMC_DIRVAR_POSITION) {
    double D[3];
    ld3(P.dir, n, i, D);
    dz = cos_to_position(D, X);
  }
  obs_vars(P.dimension, G.frame, X, dz, 15u, v);
  int64_t idx = time_index * B.time_stride;
  for (int k = 0; k < 4; ++k) {
    if (P.dimension == 2 && k == 2) continue;
    int64_t b = bin_search(B.edges[k], B.n_edges[k], v[k]);
    if (b < 0) return;
    idx += b * B.stride[k];
  }
  if (!acoustics && !P.p[i]) idx += B.mode_stride;
  atomicAdd(E + idx, 1u);
}

static int check_particles(const rtmc_particles* P) {
  RTMC_REQUIRE(P != nullptr, "particles is NULL");
  RTMC_REQUIRE(P->n >= 0, "particle count must be >= 0");
  RTMC_REQUIRE(P->dimension == 2 || P->dimension == 3, "particle dimension must be 2 or 3");
  RTMC_REQUIRE(P->n == 0 || (P->x && P->dir && P->perp && P->p && P->alive && P->t),
               "particle arrays must be device pointers");
  return RTMC_OK;
}

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace rtmc

using namespace rtmc;

extern "C" int rtmc_init_particles(const rtmc_source* source, int32_t acoustics, const rtmc_draws* draws,
                                   rtmc_particles* P, void* stream) {
  clear_error();
  int e = check_particles(P);
  if (e) return e;
  RTMC_REQUIRE(source != nullptr, "source is NULL");
  RTMC_REQUIRE(source->type == RTMC_SRC_POINT || source->type == RTMC_SRC_PLANE, "unknown source type %d",
               source->type);
  RTMC_REQUIRE(source->type == RTMC_SRC_POINT || P->dimension == 3, "plane sources are 3-D only");
  if (source->type == RTMC_SRC_POINT && source->radial.n != 0) {
    e = check_table(&source->radial, "source.radial");
    if (e) return e;
  }
  e = check_draws(draws, 8, "rtmc_init_particles");
  if (e) return e;
  if (P->n == 0) return RTMC_OK;
  SlotDraws D = to_slotdraws(draws, P->n);
  hipLaunchKernelGGL(k_init, dim3(nblocks(P->n)), dim3(kBlock), 0, (hipStream_t)stream, *source, acoustics, D, *P);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

extern "C" int rtmc_propagate_smalldt(const rtmc_material* mat, const rtmc_geometry* geometry, rtmc_particles* P,
                                      double T, double T_eff, const rtmc_draws* draws, rtmc_substep_info* info,
                                      void* stream) {
  clear_error();
  int e = check_particles(P);
  if (e) return e;
  e = check_geometry(geometry);
  if (e) return e;
  e = check_material(mat, geometry);
  if (e) return e;
  e = check_draws(draws, 6 + 2 * geometry->n_bnd, "rtmc_propagate_smalldt");
  if (e) return e;
  hipStream_t st = (hipStream_t)stream;
  double Te = T_eff;
  if (std::isnan(T_eff)) {  // :4 T = T - mean(P.t)
    if (P->n == 0) {
      Te = T;
    } else {
      Scratch s;
      e = get_scratch(&s);
      if (e) return e;
      hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, st, P->t, P->n, s.sum);
      RTMC_HIP(hipGetLastError());
      double sum = 0.0;
      RTMC_HIP(hipMemcpyAsync(&sum, s.sum, sizeof(double), hipMemcpyDeviceToHost, st));
      RTMC_HIP(hipStreamSynchronize(st));
      Te = T - sum / (double)P->n;
    }
  }
  // :5-7 sub-step
  double mftmin = mat->acoustics ? mat->mft[0] : fmin(mat->mft[0], mat->mft[1]);
  double dt = fmin(0.01 * mftmin, Te);
  int64_t nsub = (Te == 0.0) ? 0 : (int64_t)ceil(Te / dt);
  dt = nsub > 0 ? Te / (double)nsub : 0.0;
  // :31-82 intrinsic attenuation
  double QP = mat->QP, QS = mat->acoustics ? mat->QP : mat->QS;
  int useQ = std::isfinite(QP) || std::isfinite(QS);
  double pKP = 0.0, pKS = 0.0;
  if (useQ) {
    double omega = 2.0 * kPi * mat->frequency;
    pKP = std::isfinite(QP) ? 1.0 - exp(-omega * dt / QP) : 0.0;
    pKS = std::isfinite(QS) ? 1.0 - exp(-omega * dt / QS) : 0.0;
  }
  if (info) {
    info->T_eff = Te;
    info->dt = dt;
    info->n_sub = nsub;
    info->pkill_p = pKP;
    info->pkill_s = pKS;
  }
  if (draws->kind == RTMC_DRAWS_REPLAY && draws->n_steps < nsub) {
    set_error("rtmc_propagate_smalldt: replay holds %lld sub-steps, %lld needed", (long long)draws->n_steps,
              (long long)nsub);
    return RTMC_ERR_DRAWS;
  }
  if (P->n == 0) return RTMC_OK;
  MatDev M = to_matdev(mat);
  SlotDraws D = to_slotdraws(draws, P->n);
  if (P->dimension == 2)
    hipLaunchKernelGGL(k_smalldt<2>, dim3(nblocks(P->n)), dim3(kBlock), 0, st, M, *geometry, *P, dt, nsub, pKP, pKS,
                       useQ, D, Te);
  else
    hipLaunchKernelGGL(k_smalldt<3>, dim3(nblocks(P->n)), dim3(kBlock), 0, st, M, *geometry, *P, dt, nsub, pKP, pKS,
                       useQ, D, Te);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

extern "C" int rtmc_propagate_events(const rtmc_material* mat, rtmc_particles* P, double T, const rtmc_draws* draws,
                                     void* stream) {
  clear_error();
  int e = check_particles(P);
  if (e) return e;
  e = check_material(mat, nullptr);
  if (e) return e;
  e = check_draws(draws, 6, "rtmc_propagate_events");
  if (e) return e;
  if (P->n == 0) return RTMC_OK;
  hipStream_t st = (hipStream_t)stream;
  Scratch s;
  e = get_scratch(&s);
  if (e) return e;
  RTMC_HIP(hipMemsetAsync(s.flag, 0, sizeof(int), st));
  MatDev M = to_matdev(mat);
  SlotDraws D = to_slotdraws(draws, P->n);
  if (P->dimension == 2)
    hipLaunchKernelGGL(k_events<2>, dim3(nblocks(P->n)), dim3(kBlock), 0, st, M, *P, T, D, s.flag);
  else
    hipLaunchKernelGGL(k_events<3>, dim3(nblocks(P->n)), dim3(kBlock), 0, st, M, *P, T, D, s.flag);
  RTMC_HIP(hipGetLastError());
  int flag = 0;
  RTMC_HIP(hipMemcpyAsync(&flag, s.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  RTMC_HIP(hipStreamSynchronize(st));
  if (flag & 1) {
    set_error("rtmc_propagate_events: replay stream holds %lld iterations, more needed", (long long)draws->n_steps);
    return RTMC_ERR_DRAWS;
  }
  if (flag & 2) {
    set_error("rtmc_propagate_events: expected event count per interval above 700 (Poisson sampler limit)");
    return RTMC_ERR_ARG;
  }
  return RTMC_OK;
}

extern "C" int rtmc_observe(const rtmc_geometry* geometry, const rtmc_bins* bins, int32_t acoustics,
                            const rtmc_particles* P, uint32_t* E, int64_t time_index, void* stream) {
  clear_error();
  int e = check_particles(P);
  if (e) return e;
  e = check_geometry(geometry);
  if (e) return e;
  RTMC_REQUIRE(bins != nullptr && E != nullptr, "bins and E must be given");
  RTMC_REQUIRE(bins->direction_variable == RTMC_DIRVAR_DIRZ || bins->direction_variable == RTMC_DIRVAR_POSITION,
               "direction_variable must be RTMC_DIRVAR_DIRZ or RTMC_DIRVAR_POSITION");
  for (int k = 0; k < 4; ++k) {
    if (P->dimension == 2 && k == 2) continue;
    RTMC_REQUIRE(bins->n_edges[k] >= 2 && bins->edges[k] != nullptr, "variable %d needs >= 2 edges", k + 1);
  }
  if (P->n == 0) return RTMC_OK;
  hipLaunchKernelGGL(k_observe, dim3(nblocks(P->n)), dim3(kBlock), 0, (hipStream_t)stream, *geometry, *bins,
                     acoustics, *P, E, time_index);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

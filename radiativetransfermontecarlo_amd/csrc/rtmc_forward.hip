// rtmc_forward.hip -- host side of the persistent tracing kernel
// (rtmc_trace.hpp): rtmc_forward, rtmc_forward_batched,
// rtmc_adjoint_batched and rtmc_autotune_batched.
//
// Every launch is a list of problems in device memory (one for
// rtmc_forward); the kernel instances live in the rtmc_trace_*.hip units,
// one per dimension and flight precision, so that they compile in parallel.
#include "rtmc_trace.hpp"

#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <vector>

// Device scratch of rtmc_forward_noalloc: the problem record, the particle
// counters, the inverse-CDF guide and the time-major scratch of a call are
// carved from this array instead of the memory pool.  Each GPU holds its own
// copy (at its own address, found with hipGetSymbolAddress) until the
// process ends; calls on one GPU share it, so overlapping calls race.
static __device__ __align__(256) unsigned char scratch[RTMC_NOALLOC_SCRATCH_BYTES];

namespace rtmc {

// Zero-fill of n 32-bit words (the no-alloc path clears its scratch with a
// kernel rather than hipMemsetAsync).
__global__ __launch_bounds__(256) void k_fill_zero(uint32_t* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// defined in rtmc_trace_{2,3}d_{f64,f32}.hip
template <int DIM, typename R>
int launch_trace(const LaunchParams& p, const ProbDev* probs, int hm, bool quant, bool ac, bool bnd, int sink,
                 size_t smem, unsigned grid, unsigned block, hipStream_t st);

// E[base + sp + it ts + m ms] += T[(sp + m tsp) nt_pad + it - shift] for sp < tsp,
// it >= shift (shift 1: the scratch of pairs from odd times): the
// time-major scratch added into the reference layout through 64 x 64 LDS tiles
// (coalesced reads along time, coalesced writes along the bins; zero counts
// are skipped).  base: the first cell the scratch covers (a single direction
// bin, found as the tracing kernel finds it: the bin of acos(cos(pi/2)) among
// dir_edges; no dir_edges = 0).
__global__ __launch_bounds__(256) void k_tmajor_add(const uint32_t* __restrict__ T, uint32_t* __restrict__ E,
                                                    int64_t tsp, const double* __restrict__ dir_edges, int n_dir,
                                                    int64_t dir_stride, int64_t ts, int64_t ms, int nt, int nt_pad,
                                                    int shift) {
  __shared__ uint32_t tile[64][65];
  int64_t base = 0;
  if (dir_edges) {
    const int64_t k = bin_search(dir_edges, n_dir, acos(fmin(1.0, fmax(-1.0, cos(kPi / 2)))));
    if (k < 0) return;  // the tracing kernel deposited nothing in the scratch
    base = k * dir_stride;
  }
  // the tile covers scratch columns [c0, c0 + 64) of 64 bins, i.e. times c0 + shift + c;
  // rows start at even words (nt_pad is even), so columns are read in aligned pairs
  const int64_t sp0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int64_t m = blockIdx.z;
  for (int k = threadIdx.x; k < 64 * 32; k += blockDim.x) {
    const int r = k >> 5, c = (k & 31) * 2;
    const int64_t sp = sp0 + r;
    uint2 v = make_uint2(0u, 0u);
    if (sp < tsp && c0 + c < nt_pad) v = *(const uint2*)(T + (sp + m * tsp) * nt_pad + c0 + c);
    tile[r][c] = v.x;
    tile[r][c + 1] = v.y;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
    const int c = k >> 6, r = k & 63;
    const int64_t sp = sp0 + r;
    const int it = c0 + c + shift;
    const uint32_t v = tile[r][c];
    if (v && sp < tsp && it < nt) E[base + sp + (int64_t)it * ts + m * ms] += v;
  }
}

// Runs of three: E[base + sp + it ts + m ms] += sum over the n_sub scratch
// copies U_g (stride sub_words 64-bit words) and the three shifts s <= it of
// counter (it - s) % 3 of U_g[((sp + m tsp) * 3 + s) * tw + (it - s) / 3]
// (21-bit counters at bits 0, 21, 42).  A 64 x 64 tile of (bin, time) counts
// is summed in LDS in three passes, one per shift: within a pass the words of
// a row cover disjoint times, so each word of every copy is read once (along
// the row, coalesced) and added without atomics; the tile then goes out along
// the bins as in k_tmajor_add (zero counts skipped).  All copies are summed in
// one pass, so E is read and written once.
__global__ __launch_bounds__(256) void k_tmajor_add3(const unsigned long long* __restrict__ U, int n_sub,
                                                     int64_t sub_words, uint32_t* __restrict__ E, int64_t tsp,
                                                     const double* __restrict__ dir_edges, int n_dir,
                                                     int64_t dir_stride, int64_t ts, int64_t ms, int nt, int tw) {
  __shared__ uint32_t tile[64][65];
  int64_t base = 0;
  if (dir_edges) {
    const int64_t k = bin_search(dir_edges, n_dir, acos(fmin(1.0, fmax(-1.0, cos(kPi / 2)))));
    if (k < 0) return;  // the tracing kernel deposited nothing in the scratch
    base = k * dir_stride;
  }
  const int64_t sp0 = (int64_t)blockIdx.x * 64;
  const int it0 = blockIdx.y * 64;
  const int it1 = min(it0 + 64, nt);  // times [it0, it1) of this tile
  const int64_t m = blockIdx.z;
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) tile[k >> 6][k & 63] = 0u;
  for (int s = 0; s < 3; ++s) {
    __syncthreads();
    if (it1 - 1 < s) continue;  // no time of the tile is covered by this shift
    // words w with 3w + s + {0, 1, 2} meeting [it0, it1)
    const int w0 = it0 > s ? (it0 - s) / 3 : 0;
    const int w1 = (it1 - 1 - s) / 3;
    const int nw = w1 - w0 + 1;
    for (int k = threadIdx.x; k < 64 * nw; k += blockDim.x) {
      const int r = k / nw, w = w0 + (k - r * nw);
      const int64_t sp = sp0 + r;
      if (sp >= tsp) continue;
      const size_t off = ((size_t)(sp + m * tsp) * 3u + (size_t)s) * (size_t)tw + (size_t)w;
      uint32_t c0 = 0u, c1 = 0u, c2 = 0u;
      for (int g = 0; g < n_sub; ++g) {
        const unsigned long long v = U[(size_t)g * (size_t)sub_words + off];
        c0 += (uint32_t)(v & 0x1FFFFFull);
        c1 += (uint32_t)((v >> 21) & 0x1FFFFFull);
        c2 += (uint32_t)(v >> 42);
      }
      const int t = 3 * w + s;
      if (t >= it0 && t < it1) tile[r][t - it0] += c0;
      if (t + 1 >= it0 && t + 1 < it1) tile[r][t + 1 - it0] += c1;
      if (t + 2 >= it0 && t + 2 < it1) tile[r][t + 2 - it0] += c2;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
    const int c = k >> 6, r = k & 63;
    const int64_t sp = sp0 + r;
    const int it = it0 + c;
    const uint32_t v = tile[r][c];
    if (v && sp < tsp && it < nt) E[base + sp + (int64_t)it * ts + m * ms] += v;
  }
}

// Global guide of every inverse-CDF table of the launch: cell c of table t of
// problem p gets the largest knot index k <= n-2 with x[k] <= c/G (0 when
// c/G < x[0]), by a binary search per cell (Chen and Asau's guide table, here
// with about one cell per knot so that a lookup brackets one interval).
// row0 + blockIdx.y = p * ntab + t.
__global__ __launch_bounds__(256) void k_guide_build(const ProbDev* __restrict__ probs, int ntab, int log2G,
                                                     int row0) {
  const int row = row0 + (int)blockIdx.y;
  const ProbDev& P = probs[row / ntab];
  const int t = row % ntab;
  const TableDev& T = P.M.inv[t];
  const int64_t G = (int64_t)1 << log2G;
  int* g = const_cast<int*>(P.gg) + (size_t)t * (size_t)(G + 1);
  const double* __restrict__ x = T.x;
  const int64_t n = T.n;
  const double invG = 1.0 / (double)G;  // exact
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= G; c += (int64_t)gridDim.x * blockDim.x) {
    const double q = (double)c * invG;
    int64_t lo = 0, hi = n - 1;
    if (q >= x[0]) {
      while (hi - lo > 1) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (x[mid] <= q) lo = mid;
        else hi = mid;
      }
    }
    g[c] = (int)lo;
  }
}

namespace {

int getenv_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

// Launch configuration resolved from the shared part of a batch.
struct Plan {
  LaunchParams p;
  int dim;
  int hm;  // HM_*
  bool quant, ac, bnd, fp32;
  size_t smem;
  unsigned grid, block;
  int64_t hsize;  // words of one problem's histogram
  int n_modes;
  int64_t tsize;  // HM_TMAJOR: words of one problem's time-major scratch
  bool tmaj_auto; // HM_TMAJOR chosen by default (falls back to HM_GLOBAL without memory)
  bool global_guide;  // build the global guide tables (else the LDS guide, if any)
  int guide_finer;    // cell doublings of the global guide beyond one cell per knot (-1: automatic)
  int64_t run_particles;  // runs of three: particles per scratch copy (<= kRunMax)
  int64_t sub_words;      // runs of three: 64-bit words of one scratch copy
};

// a 21-bit counter of the runs-of-three scratch takes at most one count per particle
constexpr int64_t kRunMax = ((int64_t)1 << 21) - 1;

int check_shared(const rtmc_geometry& g, const rtmc_bins& B, int n_times, int time_steps, int precision) {
  int e = check_geometry(&g);
  if (e) return e;
  RTMC_REQUIRE(n_times >= 1, "n_times must be >= 1");
  RTMC_REQUIRE(time_steps == 0 || time_steps == 1, "material.timeSteps must be 0 or 1");
  RTMC_REQUIRE(precision == RTMC_PREC_FP64 || precision == RTMC_PREC_FP32_FLIGHT,
               "precision must be RTMC_PREC_FP64 or RTMC_PREC_FP32_FLIGHT");
  if (time_steps == 1) {
    // propagateParticle has neither Q nor boundaries (propagateParticle.m:1-81)
    RTMC_REQUIRE(g.n_bnd == 0, "timeSteps = 1 (propagateParticle) does not handle boundaries");
  }
  for (int v = 0; v < 4; ++v) {
    if (g.dimension == 2 && v == 2) continue;
    RTMC_REQUIRE(B.n_edges[v] >= 2 && B.edges[v] != nullptr, "variable %d needs >= 2 edges", v + 1);
  }
  RTMC_REQUIRE(B.direction_variable == RTMC_DIRVAR_DIRZ || B.direction_variable == RTMC_DIRVAR_POSITION,
               "direction_variable must be RTMC_DIRVAR_DIRZ or RTMC_DIRVAR_POSITION");
  return RTMC_OK;
}

int check_item(const rtmc_material& m, const rtmc_source& s, int64_t n, const rtmc_geometry& g) {
  int e = check_material(&m, &g);
  if (e) return e;
  RTMC_REQUIRE(n >= 0, "n_particles must be >= 0");
  RTMC_REQUIRE(s.type == RTMC_SRC_POINT || g.dimension == 3, "plane sources are 3-D only");
  if (s.type == RTMC_SRC_POINT && s.radial.n != 0) {
    e = check_table(&s.radial, "source.radial");
    if (e) return e;
  }
  return RTMC_OK;
}

// Per-problem record: Philox key, sub-step of propagateParticleSmallDt.m:4-7
// with T - mean(P.t) = dt, per-sub-step scattering and removal rates.
int make_prob(const rtmc_material& m, const rtmc_source& s, int64_t n, uint64_t offset, uint64_t seed, double dt,
              int n_times, int time_steps, uint32_t* E, uint64_t* counters, ProbDev* P) {
  memset((void*)P, 0, sizeof(*P));
  P->M = to_matdev(&m);
  P->S = s;
  P->n = n;
  P->pid0 = offset;
  uint64_t key = seed ^ 0x52544D4346574400ull;  // forward-kernel stream family
  P->k0 = (uint32_t)key;
  P->k1 = (uint32_t)(key >> 32);
  double mftP = m.mft[0], mftS = m.acoustics ? m.mft[0] : m.mft[1];
  double mftmin = m.acoustics ? mftP : fmin(mftP, mftS);
  double dts = fmin(0.01 * mftmin, dt);
  int64_t nsub = (dt > 0.0) ? (int64_t)ceil(dt / dts) : 1;
  RTMC_REQUIRE(nsub >= 1 && nsub < ((int64_t)1 << 40), "sub-step count out of range");
  dts = dt / (double)nsub;
  P->nsub = nsub;
  P->dt_sub = dts;
  P->lamS[0] = dts / mftP;
  P->lamS[1] = dts / mftS;
  double QP = m.QP, QS = m.acoustics ? m.QP : m.QS;
  P->useQ = (time_steps == 0) && (std::isfinite(QP) || std::isfinite(QS));
  double omega = 2.0 * kPi * m.frequency;
  P->lamK[0] = (P->useQ && std::isfinite(QP)) ? omega * dts / QP : 0.0;
  P->lamK[1] = (P->useQ && std::isfinite(QS)) ? omega * dts / QS : 0.0;
  for (int k = 0; k < 2; ++k) {
    P->ilamS[k] = P->lamS[k] > 0.0 ? 1.0 / P->lamS[k] : 0.0;
    P->ilamK[k] = P->lamK[k] > 0.0 ? 1.0 / P->lamK[k] : 0.0;
  }
  // scheduling weight: an event costs about six deposits
  double T = dt * (double)(n_times - 1);
  double ev = std::isfinite(mftmin) && mftmin > 0.0 ? T / mftmin : 0.0;
  P->cost = (double)(n_times - 1) + 6.0 * std::min(ev, 1e9);
  P->E = E;
  P->counters = (unsigned long long*)counters;
  return RTMC_OK;
}

// Shared setup: edges, LDS budget, histogram placement, grid.
// ev_per_obs: events (scatterings and removals) per observation interval of a
// particle, weighted over the launch's particles; it sets the default event batch.
// lt_cap: knots of the largest compressed inverse-CDF table of the launch (0: the
// problems draw from their full tables in global memory).
int make_plan(const rtmc_geometry& g, const rtmc_bins& B, int n_times, double dt, int time_steps, int precision,
              int acoustics, int block_size, int ppt, int hist_mode, int event_batch, int64_t n_total, int n_prob,
              bool adjoint, double ev_per_obs, int lt_cap, Plan* out) {
  Plan& pl = *out;
  LaunchParams& p = pl.p;
  memset((void*)&p, 0, sizeof(p));
  const int dim = g.dimension;
  p.G = g;
  int total = 0;
  p.present = 0;
  for (int v = 0; v < 4; ++v) {
    p.edges[v] = B.edges[v];
    p.ne[v] = B.n_edges[v];
    RTMC_REQUIRE(B.stride[v] >= 0 && B.stride[v] < ((int64_t)1 << 31), "bin stride %d out of range", v + 1);
    p.stride[v] = (uint32_t)B.stride[v];
    p.eoff[v] = total;
    if (dim == 2 && v == 2) continue;
    p.present |= 1u << v;
    total += B.n_edges[v];
  }
  RTMC_REQUIRE(B.time_stride >= 1 && B.time_stride < ((int64_t)1 << 31),
               "time_stride must be in [1, 2^31): one time slice of at most 2^31 - 1 bins");
  p.time_stride = B.time_stride;
  p.time_stride32 = (uint32_t)B.time_stride;
  p.mode_stride = B.mode_stride;
  p.dirvar = B.direction_variable;
  p.n_times = n_times;
  p.dt_obs = dt;
  p.n_prob = n_prob;
  // histogram extent: all time slices and modes
  int64_t hsize = B.time_stride * n_times;
  if (!acoustics) hsize = std::max(hsize, 2 * B.mode_stride);
  pl.hsize = hsize;

  int dev = 0;
  RTMC_HIP(hipGetDevice(&dev));
  int ncu = 0, max_smem = 0;
  RTMC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  RTMC_HIP(hipDeviceGetAttribute(&max_smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  // what the plan may size: the work-group's LDS less the kernel's static part
  max_smem -= kTraceStaticLds;

  size_t edge_bytes = (size_t)total * sizeof(double);
  bool edges_lds = edge_bytes <= 32 * 1024;
  p.edges_total = edges_lds ? total : 0;
  if (!edges_lds) edge_bytes = 0;
  size_t hist_bytes = (size_t)hsize * sizeof(uint32_t);  // uint32 counts or float weights
  p.ntab = acoustics ? 1 : 4;
  int guide = 2048;
  if (const char* gc = getenv("RTMC_GUIDE_CELLS")) guide = atoi(gc);
  RTMC_REQUIRE(guide >= 0 && guide <= (1 << 16) && (guide & (guide - 1)) == 0,
               "RTMC_GUIDE_CELLS must be 0 or a power of two <= 65536");
  auto guide_bytes = [&](int G) { return G > 0 ? (size_t)p.ntab * (size_t)(G + 1) * sizeof(int) : (size_t)0; };
  // compressed tables in LDS: per table lt_cap 16-byte knots and a guide of
  // G + 1 uint16 (G the power of two >= lt_cap, in [64, 4096]), padded to 16 bytes
  p.lt_cap = lt_cap;
  p.lt_glog2 = 6;
  while (p.lt_glog2 < 12 && (1 << p.lt_glog2) < lt_cap) ++p.lt_glog2;
  p.lt_gstride = (((1 << p.lt_glog2) + 1) + 7) & ~7;
  const size_t lt_bytes =
      lt_cap > 0 ? (size_t)p.ntab * ((size_t)lt_cap * sizeof(rtmc_lds_knot) + (size_t)p.lt_gstride * 2) : 0;
  RTMC_REQUIRE(hist_mode >= 0 && hist_mode <= 3, "hist_mode must be 0, 1, 2 or 3");
  bool ldsh;
  if (hist_mode == 1 || hist_mode == 3) ldsh = false;
  else if (hist_mode == 2) ldsh = true;
  else ldsh = (edge_bytes + hist_bytes + lt_bytes) <= (size_t)max_smem && hsize <= (int64_t)1 << 20;
  pl.n_modes = acoustics ? 1 : 2;
  // time-major scratch: explicit (3), or by default whenever the histogram is in global memory
  bool tmaj = !adjoint && !ldsh && n_times >= 2 && (hist_mode == 3 || hist_mode == 0);
  p.nt_pad = n_times + (n_times & 1);
  // 2-D, dir_z binned as the outermost variable: the scratch keeps one direction bin
  const int64_t ndir = (int64_t)B.n_edges[3] - 1;
  p.tcomp = tmaj && dim == 2 && B.direction_variable == RTMC_DIRVAR_DIRZ && ndir > 1 && B.stride[3] > 0 &&
            B.stride[3] * ndir == B.time_stride && getenv_int("RTMC_TMAJOR_ONE_DIRECTION", 1) != 0;
  p.tsp = (uint32_t)(p.tcomp ? B.stride[3] : B.time_stride);
  pl.tsize = tmaj ? (int64_t)p.tsp * pl.n_modes * (int64_t)p.nt_pad : 0;
  // a second scratch shifted by one time takes the pairs that start at odd times
  // when the two fit in a quarter of the device's memory (1e9-particle case, two
  // 28.4 GB scratches: step 2025 -> 1946 ms, with the compressed angle tables
  // 1660 -> 1604 ms; run_plan drops it when the allocation fails)
  {
    const int op = getenv_int("RTMC_ODD_PAIRS", 1);  // 0 off, 1 when they fit, 2 always
    size_t mem_free = 0, mem_total = 0;
    if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) {
      (void)hipGetLastError();
      mem_total = 0;
    }
    const int64_t lim = std::max<int64_t>((int64_t)2 << 30, (int64_t)(mem_total / 8));
    p.odd_pairs = tmaj && (op == 2 || (op == 1 && pl.tsize * 4 <= lim));
  }
  // runs of three (21-bit counters, one scratch copy per run_particles
  // particles) when all the copies fit in 4 GiB
  {
    const int r3 = getenv_int("RTMC_RUNS3", 1);  // 0 off, 1 when <= 4 GiB, 2 always
    const int64_t rp = getenv_int("RTMC_RUN_PARTICLES", (int)kRunMax);
    RTMC_REQUIRE(rp >= 1 && rp <= kRunMax, "RTMC_RUN_PARTICLES must be in [1, %lld]", (long long)kRunMax);
    pl.run_particles = rp;
    p.tw = (n_times + 2) / 3;
    const int64_t rows = (int64_t)p.tsp * pl.n_modes;
    pl.sub_words = rows * 3 * (int64_t)p.tw;
    const int64_t nsub = n_total / rp + n_prob;  // scratch copies, at most
    p.runs3 = tmaj && B.direction_variable == RTMC_DIRVAR_DIRZ && (int64_t)rows * p.nt_pad < ((int64_t)1 << 32) &&
              (r3 == 2 || (r3 == 1 && nsub * pl.sub_words * 8 <= ((int64_t)4 << 30)));
    if (p.runs3) p.odd_pairs = 0;
    p.run_n = (int)rp;
    // 1: every run as a 64-bit atomic (a run of one otherwise goes out as a 32-bit
    // atomic on the low half of its word)
    p.run64 = getenv_int("RTMC_RUNS3_ATOMIC64", 0) != 0;
    p.run_words = pl.sub_words;
  }
  pl.hm = ldsh ? HM_LDS : (tmaj ? HM_TMAJOR : HM_GLOBAL);
  pl.tmaj_auto = tmaj && hist_mode == 0;
  if (ldsh && edge_bytes + hist_bytes + lt_bytes > (size_t)max_smem) {
    set_error("LDS %s of %zu bytes (with %zu bytes of bin edges and %zu of compressed angle tables) exceeds the %d "
              "bytes of LDS a work-group has beside the kernel's own", adjoint ? "weight table" : "histogram",
              hist_bytes, edge_bytes, lt_bytes, max_smem);
    return RTMC_ERR_ARG;
  }
  RTMC_REQUIRE(edge_bytes + lt_bytes <= (size_t)max_smem,
               "compressed angle tables of %zu bytes (with %zu bytes of bin edges) exceed the %d bytes of LDS a "
               "work-group has beside the kernel's own: compress them with fewer knots",
               lt_bytes, edge_bytes, max_smem);
  // The global guide (run_plan) replaces the LDS guide, which then takes no LDS; if
  // the global guide cannot be allocated the lookups search the whole table.
  // RTMC_GLOBAL_GUIDE=0 keeps the LDS guide, which takes what the histogram leaves.
  // Compressed tables in LDS need neither.
  pl.global_guide = getenv_int("RTMC_GLOBAL_GUIDE", 1) != 0 && lt_cap == 0;
  pl.guide_finer = getenv_int("RTMC_GUIDE_FINER", -1);
  RTMC_REQUIRE(pl.guide_finer >= -1 && pl.guide_finer <= 6, "RTMC_GUIDE_FINER must be -1 (automatic) or in [0, 6]");
  if (pl.global_guide || lt_cap > 0) guide = 0;
  while (guide > 0 && edge_bytes + guide_bytes(guide) + (ldsh ? hist_bytes : 0) > (size_t)max_smem) guide >>= 1;
  if (guide < 64) guide = 0;
  // time-major mode: a cache of hot counters, 8 B a slot: 4096 slots (32 KB).  FP32
  // flight had used 2048 while its kernel was bound by its own work, as its
  // instances use few registers and lose half their work-groups per CU to the
  // larger cache (1e9-particle case: 1.96 s with 2048, 2.19 s with 4096); since
  // the polar angle's sin/cos went through sincospi that kernel waits on the
  // atomics instead, and the requests the larger cache saves pay (1637 -> 1626 ms,
  // means of four runs each in ABBA order).  FP64 3-D half-space 381.6 -> 375.8 ms with
  // 4096, its occupancy being set by registers.  The LDS guide, if any, shrinks to
  // keep the work-group within 40 KB.
  // 3-D pairs with the full tables (`wide`): work-groups of 1024 threads, one per
  // CU (16 wavefronts, what registers or LDS allowed before too), sharing one
  // cache of 8192 slots instead of four work-groups with 4096 each; the
  // 1e9-particle kernel 1636 -> 1618 ms (1621 with 512 threads, 1631 with 512
  // threads and 4096 slots), with FP32 flight 1628 -> 1609 ms, ABBA order.
  // 2-D runs of three too: their particles share a plane, so more of their
  // counters are hot (2-D elastic kernel 40.96 -> 37.58 ms, ABBA order).
  const bool wide = tmaj && lt_cap == 0 && (dim == 3) != (p.runs3 != 0) && (block_size == 0 || block_size >= 512);
  int hot = 0;
  if (tmaj) {
    hot = wide ? 8192 : 4096;
    if (const char* h = getenv("RTMC_HOT_SLOTS")) hot = atoi(h);
    RTMC_REQUIRE(hot == 0 || (hot >= 2 && hot <= (1 << 13) && (hot & (hot - 1)) == 0),
                 "RTMC_HOT_SLOTS must be 0 or a power of two in [2, 8192]");
    const size_t budget = 40 * 1024;
    while (guide > 256 && edge_bytes + guide_bytes(guide) + 8 + 8 * (size_t)hot > budget) guide >>= 1;
    // The compressed tables do not shrink the cache: their instances hold more
    // registers (3 waves per SIMD for FP64 flight), which leaves them the LDS of 3
    // work-groups per CU (1e9-particle case, 450-knot table: 1639 ms with 4096 slots,
    // 1744 ms with 2048).
  }
  p.guide_n = guide;
  p.hist_size = ldsh ? hsize : 0;
  pl.smem = edge_bytes + guide_bytes(guide) + (ldsh ? hist_bytes : 0);
  p.lt_off = (int)((pl.smem + 15) & ~(size_t)15);
  p.lt_goff = p.lt_off + (int)((size_t)p.ntab * (size_t)lt_cap * sizeof(rtmc_lds_knot));
  if (lt_cap > 0) pl.smem = (size_t)p.lt_off + lt_bytes;
  p.hot_slots = hot;
  p.hot_times = 1 << 30;
  if (const char* h = getenv("RTMC_HOT_TIMES")) p.hot_times = atoi(h);
  {
    const int64_t cap = (1 << 24) - 65536;  // a slot's count has 24 bits
    const int64_t m = getenv_int("RTMC_HOT_MAX_PARTICLES", (int)cap);
    RTMC_REQUIRE(m >= 256 && m <= cap, "RTMC_HOT_MAX_PARTICLES must be in [256, %lld]", (long long)cap);
    p.hot_max_taken = (unsigned)m;
  }
  p.hot_shift = 32;
  for (int v = hot; v > 1; v >>= 1) p.hot_shift -= 1;
  p.hot_off = (int)((pl.smem + 7) & ~(size_t)7);
  if (hot) pl.smem = (size_t)p.hot_off + 8 * (size_t)hot;
  RTMC_REQUIRE(pl.smem <= (size_t)max_smem, "%zu bytes of LDS per work-group exceed the %d available beside the "
               "kernel's own", pl.smem, max_smem);

  unsigned block = block_size > 0 ? (unsigned)block_size : ((ldsh || wide) ? 1024u : 256u);
  RTMC_REQUIRE(block % 64 == 0 && block <= ((ldsh || (tmaj && lt_cap == 0 && (dim == 3) != (p.runs3 != 0))) ? 1024u : 256u),
               "block_size must be a multiple of 64 and <= 256 (<= 1024 with the LDS histogram, or with 3-D "
               "time-major pairs and the full angle tables)");
  // resident work-groups per CU as LDS allows (the whole work-group's LDS, static part included)
  const int64_t lds_cu = (int64_t)max_smem + kTraceStaticLds, lds_wg = (int64_t)pl.smem + kTraceStaticLds;
  int64_t per_cu;
  if (ldsh) per_cu = std::max<int64_t>(1, lds_cu / lds_wg);
  else per_cu = std::max<int64_t>(1, std::min<int64_t>(2048 / block, lds_cu / lds_wg));
  RTMC_REQUIRE(ppt >= 0, "particles_per_thread must be >= 0");
  int64_t want = ppt > 0 ? (n_total + (int64_t)block * ppt - 1) / ((int64_t)block * ppt)
                         : (n_total + block - 1) / block;
  int64_t cap = (int64_t)ncu * per_cu;
  pl.grid = (unsigned)std::max<int64_t>(1, std::min(want, cap));
  pl.block = block;

  p.chunk = 256;
  // Lanes waiting for an event before a wavefront takes an event step: more when
  // events are frequent against observations (an event step then serves more
  // lanes; deposit steps lose the waiting lanes).  Measured optima (event steps
  // with block-aligned draws and deferred births): 16 at 0.01 events per
  // observation (3-D README case: 1.46 ms at 24, 1.36 at 16), 32 at 0.05 (2-D
  // elastic and half-space cases), 40 at 0.35 (lc = 0.05, f = 40, Q = 300:
  // 193.7 ms per 1e8 particles at 48, 190.3 at 40, 207.7 at 56).
  // FP32 flight runs twice the wavefronts per SIMD and prefers 48 there (1e9-particle
  // case per 1e8 particles: 185 ms at 48 against 196 at 40).  With the finer global
  // guide the FP64 full-table lookups of the time-major kernel prefer larger event
  // steps, whose many lanes overlap their dependent reads: 1e9-particle kernel
  // 1816 ms at 44, 1789 at 48, 1764 at 50, 1648 at 52, 1686 at 54, 1914 at 60; with
  // the tables in LDS 40 stays best (1536 ms; 1597 at 48).  After the polar angle's
  // sin/cos moved to sincospi, one box, two runs each: 1731 at 50, 1728 at 51, 1650
  // at 52, 1638 at 53, 1661 at 54 (and 1454 at 40, 1449 at 36, 1459 at 32 with the
  // LDS tables); 52 against 53 in ABBA order, four runs each: 1644 and 1636 ms, 53
  // losing less where the scratch's placement takes fewer atomics per second (1640
  // against 1671 ms).  The compressed-table instances take 36 (1449 against 1455 ms
  // at 40, ABBA order, four runs each).
  const bool f32 = precision == RTMC_PREC_FP32_FLIGHT;
  const int heavy = f32 ? 48 : (lt_cap > 0 ? 36 : (pl.hm == HM_TMAJOR ? 53 : 40));
  p.batch = event_batch > 0 ? event_batch : (ev_per_obs >= 0.15 ? heavy : (ev_per_obs >= 0.02 ? 32 : 16));
  if (event_batch == 0) {
    if (const char* b = getenv("RTMC_EVENT_BATCH")) p.batch = atoi(b);
  }
  RTMC_REQUIRE(p.batch >= 1 && p.batch <= 64, "event batch must be in [1, 64]");
  // Births wait for birth_min lanes when events are frequent (the birth path then
  // runs for several lanes at once): 1e9-particle case (0.35 events per
  // observation) 270 -> 236 ms per 1e8 particles with 8, the 2-D elastic and
  // half-space cases (0.05) 1-2 % faster; the README case (0.01, few events,
  // many births per event step) 7 % slower, so it keeps 1.  FP32 flight, with
  // twice the wavefronts per SIMD, loses with 8 (202.6 ms against 185.8 with 1)
  // and gains a little with 2 at its batch of 48 (180.5 ms).
  p.birth_min = f32 ? (ev_per_obs >= 0.15 ? 2 : 1) : (ev_per_obs >= 0.02 ? 8 : 1);
  if (const char* b = getenv("RTMC_BIRTH_MIN")) p.birth_min = atoi(b);
  RTMC_REQUIRE(p.birth_min >= 1 && p.birth_min <= 64, "RTMC_BIRTH_MIN must be in [1, 64]");
  pl.dim = dim;
  pl.quant = time_steps == 0;
  pl.ac = acoustics != 0;
  pl.bnd = g.n_bnd > 0;
  pl.fp32 = precision == RTMC_PREC_FP32_FLIGHT;
  return RTMC_OK;
}

// Copies the problems to the device, clears the counters, launches, frees.
// Optional timing of the tracing kernel launches (rtmc_set_trace_timing):
// a per-thread ring of event pairs recorded on the launch stream.
struct TraceTimer {
  bool on = false;
  int n = 0;
  std::vector<hipEvent_t> ev;  // 2 per slot
};
thread_local TraceTimer g_timer;
constexpr int kTimerSlots = 1024;

void timer_mark(hipStream_t st, int which) {
  TraceTimer& T = g_timer;
  if (!T.on || T.n >= kTimerSlots) return;
  const size_t k = 2 * (size_t)T.n + (size_t)which;
  while (T.ev.size() <= k) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    T.ev.push_back(e);
  }
  (void)hipEventRecord(T.ev[k], st);
  if (which == 1) T.n += 1;
}

// The stream-ordered pool keeps what it has handed out (the time-major
// scratch is GBs and is needed again by the next call).
int keep_pool_memory() {
  static thread_local int done_dev = -1;
  int dev = 0;
  RTMC_HIP(hipGetDevice(&dev));
  if (done_dev == dev) return RTMC_OK;
  hipMemPool_t pool;
  RTMC_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t thr = UINT64_MAX;
  RTMC_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
  done_dev = dev;
  return RTMC_OK;
}

// Device memory of one launch: from the stream-ordered pool (base null), or
// carved from the no-alloc scratch array (base = its device address).
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  size_t avail() const {
    const size_t off = (used + 255) & ~(size_t)255;
    return off < cap ? cap - off : 0;
  }
};

hipError_t arena_alloc(Arena* a, void** p, size_t bytes, hipStream_t st) {
  if (!a->base) return hipMallocAsync(p, bytes, st);
  const size_t off = (a->used + 255) & ~(size_t)255;
  if (off > a->cap || bytes > a->cap - off) return hipErrorOutOfMemory;
  *p = a->base + off;
  a->used = off + bytes;
  return hipSuccess;
}

hipError_t arena_free(Arena* a, void* p, hipStream_t st) { return a->base ? hipSuccess : hipFreeAsync(p, st); }

// clears bytes (a multiple of 4) at p: hipMemsetAsync, or a fill kernel in the scratch array
hipError_t arena_zero(Arena* a, void* p, size_t bytes, hipStream_t st) {
  if (!a->base) return hipMemsetAsync(p, 0, bytes, st);
  const int64_t n = (int64_t)(bytes / 4);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, st, (uint32_t*)p, n);
  return hipGetLastError();
}

// Scratch of the time-major placement per problem (32-bit words at toff[i]): T
// (and T2), or with runs of three one copy of 64-bit words per run_particles particles.
size_t tmajor_words(const Plan& pl, const std::vector<ProbDev>& probs, std::vector<int>* nsub,
                    std::vector<size_t>* toff) {
  const int np = (int)probs.size();
  nsub->assign(np, 1);
  toff->assign(np + 1, 0);
  for (int i = 0; i < np; ++i) {
    if (pl.p.runs3) (*nsub)[i] = (int)((probs[i].n + pl.run_particles - 1) / pl.run_particles);
    (*toff)[i + 1] = (*toff)[i] + (pl.p.runs3 ? 2 * (size_t)pl.sub_words * (size_t)(*nsub)[i]
                                              : (size_t)pl.tsize * (pl.p.odd_pairs ? 2 : 1));
  }
  return (*toff)[np];
}

// the time-major placement is dropped: direct global atomics need no scratch
// The same threads in work-groups of 256, for instances that take no more (the
// global-atomics ones, and the 2-D pairs that replace runs of three).
void narrow_block(Plan& pl) {
  if (pl.block > 256) {
    pl.grid *= pl.block / 256;
    pl.block = 256;
  }
}

void drop_tmajor(Plan& pl) {
  narrow_block(pl);
  pl.hm = HM_GLOBAL;
  pl.p.hot_slots = 0;
  pl.p.tcomp = 0;
  pl.p.odd_pairs = 0;
  pl.p.runs3 = 0;
}

int run_plan(Plan& pl, std::vector<ProbDev>& probs, int sink, hipStream_t st, Arena* arena = nullptr) {
  Arena pool;  // default: the stream-ordered pool
  Arena& ar = arena ? *arena : pool;
  const int np = (int)probs.size();
  const size_t pbytes = sizeof(ProbDev) * (size_t)np;
  const size_t wbytes = sizeof(unsigned long long) * (size_t)np;
  const size_t abytes = sizeof(unsigned int) * (size_t)np;
  void* buf = nullptr;
  void* tbuf = nullptr;
  void* gbuf = nullptr;
  if (!ar.base) {
    int e = keep_pool_memory();
    if (e) return e;
  }
  // problem records and counters first: nothing runs without them
  if (arena_alloc(&ar, &buf, pbytes + wbytes + abytes, st) != hipSuccess) {
    (void)hipGetLastError();
    set_error("%s of %zu bytes of problem data failed", ar.base ? "no-alloc scratch" : "hipMallocAsync",
              pbytes + wbytes + abytes);
    return RTMC_ERR_NOMEM;
  }
  // global guides of the inverse-CDF tables: about one cell per knot, G a power of two
  int64_t nmax = 2;
  for (int i = 0; i < np; ++i)
    for (int t = 0; t < pl.p.ntab; ++t) nmax = std::max<int64_t>(nmax, probs[i].M.inv[t].n);
  int lg = 10;
  while (lg < 22 && ((int64_t)1 << lg) < nmax) ++lg;
  {
    // More cells per knot when many particles share a table: a lookup then
    // brackets fewer knots (each search step is a dependent global read), for the
    // price of building and caching a larger guide.  Automatic (-1): one doubling
    // per factor two of particles per problem beyond 64 per knot, at most 16 cells
    // per knot; RTMC_GUIDE_FINER = k doubles k times.  The guides stay within 1 GiB
    // (half the free no-alloc scratch).
    int64_t parts = 0;
    for (int i = 0; i < np; ++i) parts += probs[i].n;
    parts /= std::max(1, np);
    int extra = pl.guide_finer;
    if (extra < 0) {
      extra = 0;
      while (extra < 4 && parts >= ((int64_t)64 << (extra + 1)) * nmax) ++extra;
    }
    const size_t cap = ar.base ? ar.avail() / 2 : ((size_t)1 << 30);
    while (extra > 0 && sizeof(int) * (((size_t)1 << (lg + extra)) + 1) * (size_t)pl.p.ntab * (size_t)np > cap) --extra;
    lg += extra;
  }
  const bool want_gg = pl.global_guide && pl.p.lt_cap == 0 && nmax <= ((int64_t)1 << 31) - 2;
  const size_t gcells = ((size_t)1 << lg) + 1;
  const size_t gbytes = sizeof(int) * gcells * (size_t)pl.p.ntab * (size_t)np;
  pl.p.gg_log2 = 0;
  const int lds_guide = pl.p.guide_n;  // restored below: the plan may be run again
  if (want_gg) {
    if (arena_alloc(&ar, &gbuf, gbytes, st) == hipSuccess) {
      for (int i = 0; i < np; ++i) probs[i].gg = (const int*)gbuf + gcells * (size_t)pl.p.ntab * (size_t)i;
      pl.p.gg_log2 = lg;
      pl.p.guide_n = 0;  // the LDS guide is not needed
    } else {
      gbuf = nullptr;
      (void)hipGetLastError();  // fall back to the LDS guide
    }
  }
  std::vector<int> nsub;
  std::vector<size_t> toff;
  size_t tbytes = 0;
  if (pl.hm == HM_TMAJOR) {
    tbytes = sizeof(uint32_t) * tmajor_words(pl, probs, &nsub, &toff);
    if (ar.base && tbytes > ar.avail()) {
      // the scratch array is too small for runs of three: pairs, then pairs from even times only
      if (pl.p.runs3) {
        pl.p.runs3 = 0;
        if (pl.dim == 2) narrow_block(pl);
        pl.p.odd_pairs = 1;
        tbytes = sizeof(uint32_t) * tmajor_words(pl, probs, &nsub, &toff);
      }
      if (tbytes > ar.avail() && pl.p.odd_pairs) {
        pl.p.odd_pairs = 0;
        tbytes = sizeof(uint32_t) * tmajor_words(pl, probs, &nsub, &toff);
      }
    }
    hipError_t ae = arena_alloc(&ar, &tbuf, tbytes, st);
    if (ae != hipSuccess && !pl.p.runs3 && pl.p.odd_pairs) {
      // without the second scratch of pairs from odd times
      (void)hipGetLastError();
      pl.p.odd_pairs = 0;
      tbytes = sizeof(uint32_t) * tmajor_words(pl, probs, &nsub, &toff);
      ae = arena_alloc(&ar, &tbuf, tbytes, st);
    }
    if (ae != hipSuccess) {
      tbuf = nullptr;
      (void)hipGetLastError();
      if (!pl.tmaj_auto) {
        pl.p.guide_n = lds_guide;
        pl.p.gg_log2 = 0;
        if (gbuf) (void)arena_free(&ar, gbuf, st);
        (void)arena_free(&ar, buf, st);
        if (ar.base) {
          set_error("time-major scratch of %zu bytes exceeds the %llu bytes of the no-alloc scratch (use hist_mode 0 "
                    "or 1)", tbytes, (unsigned long long)RTMC_NOALLOC_SCRATCH_BYTES);
          return RTMC_ERR_ARG;
        }
        set_error("time-major scratch of %zu bytes: %s (use hist_mode 1)", tbytes, hipGetErrorString(ae));
        return RTMC_ERR_NOMEM;
      }
      drop_tmajor(pl);
    } else {
      for (int i = 0; i < np; ++i) {
        probs[i].T = (uint32_t*)tbuf + toff[i];
        probs[i].T2 = pl.p.odd_pairs ? probs[i].T + pl.tsize : nullptr;
      }
    }
  }
  char* b = (char*)buf;
  ProbDev* dprobs = (ProbDev*)b;
  pl.p.work = (unsigned long long*)(b + pbytes);
  pl.p.active = (unsigned int*)(b + pbytes + wbytes);
  int rc = RTMC_OK;
  hipError_t e = hipMemcpyAsync(dprobs, probs.data(), pbytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = arena_zero(&ar, pl.p.work, wbytes + abytes, st);
  if (e == hipSuccess && tbuf) e = arena_zero(&ar, tbuf, tbytes, st);
  if (e != hipSuccess) {
    set_error("problem upload failed: %s", hipGetErrorString(e));
    rc = RTMC_ERR_HIP;
  }
  if (rc == RTMC_OK && gbuf) {
    const unsigned gx = (unsigned)std::min<size_t>((gcells + 255) / 256, 1024);
    const int rows = np * pl.p.ntab;
    for (int r0 = 0; r0 < rows && rc == RTMC_OK; r0 += 65535) {  // grid.y limit
      const int nr = std::min(65535, rows - r0);
      hipLaunchKernelGGL(k_guide_build, dim3(gx, (unsigned)nr), dim3(256), 0, st, (const ProbDev*)dprobs,
                         pl.p.ntab, lg, r0);
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) {
        set_error("k_guide_build launch failed: %s", hipGetErrorString(le));
        rc = RTMC_ERR_HIP;
      }
    }
  }
  if (rc == RTMC_OK) {
    timer_mark(st, 0);
    const int hm = (pl.hm == HM_TMAJOR && pl.p.runs3) ? HM_TMAJOR3 : pl.hm;
    if (pl.dim == 2) {
      rc = pl.fp32 ? launch_trace<2, float>(pl.p, dprobs, hm, pl.quant, pl.ac, pl.bnd, sink, pl.smem, pl.grid,
                                            pl.block, st)
                   : launch_trace<2, double>(pl.p, dprobs, hm, pl.quant, pl.ac, pl.bnd, sink, pl.smem, pl.grid,
                                             pl.block, st);
    } else {
      rc = pl.fp32 ? launch_trace<3, float>(pl.p, dprobs, hm, pl.quant, pl.ac, pl.bnd, sink, pl.smem, pl.grid,
                                            pl.block, st)
                   : launch_trace<3, double>(pl.p, dprobs, hm, pl.quant, pl.ac, pl.bnd, sink, pl.smem, pl.grid,
                                             pl.block, st);
    }
  }
  if (rc == RTMC_OK) timer_mark(st, 1);
  if (rc == RTMC_OK && tbuf && pl.p.runs3) {
    const int64_t ts = pl.p.time_stride;
    dim3 grid((unsigned)(((int64_t)pl.p.tsp + 63) / 64), (unsigned)((pl.p.n_times + 63) / 64), (unsigned)pl.n_modes);
    for (int i = 0; i < np && rc == RTMC_OK; ++i) {
      hipLaunchKernelGGL(k_tmajor_add3, grid, dim3(256), 0, st, (const unsigned long long*)probs[i].T, nsub[i],
                         pl.sub_words, probs[i].E, (int64_t)pl.p.tsp,
                         pl.p.tcomp ? pl.p.edges[3] : (const double*)nullptr, pl.p.ne[3], (int64_t)pl.p.stride[3], ts,
                         pl.p.mode_stride, pl.p.n_times, pl.p.tw);
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) {
        set_error("k_tmajor_add3 launch failed: %s", hipGetErrorString(le));
        rc = RTMC_ERR_HIP;
      }
    }
  } else if (rc == RTMC_OK && tbuf) {
    const int64_t ts = pl.p.time_stride;
    dim3 grid((unsigned)(((int64_t)pl.p.tsp + 63) / 64), (unsigned)((pl.p.nt_pad + 63) / 64), (unsigned)pl.n_modes);
    for (int i = 0; i < np && rc == RTMC_OK; ++i) {
      for (int sh = 0; sh <= (pl.p.odd_pairs ? 1 : 0); ++sh) {
        hipLaunchKernelGGL(k_tmajor_add, grid, dim3(256), 0, st, (const uint32_t*)(sh ? probs[i].T2 : probs[i].T),
                           probs[i].E, (int64_t)pl.p.tsp, pl.p.tcomp ? pl.p.edges[3] : (const double*)nullptr,
                           pl.p.ne[3], (int64_t)pl.p.stride[3], ts, pl.p.mode_stride, pl.p.n_times, pl.p.nt_pad,
                           sh);
      }
      hipError_t le = hipGetLastError();
      if (le != hipSuccess) {
        set_error("k_tmajor_add launch failed: %s", hipGetErrorString(le));
        rc = RTMC_ERR_HIP;
      }
    }
  }
  pl.p.guide_n = lds_guide;
  pl.p.gg_log2 = 0;
  hipError_t fe = arena_free(&ar, buf, st);
  hipError_t fe2 = tbuf ? arena_free(&ar, tbuf, st) : hipSuccess;
  hipError_t fe3 = gbuf ? arena_free(&ar, gbuf, st) : hipSuccess;
  if (rc == RTMC_OK && (fe != hipSuccess || fe2 != hipSuccess || fe3 != hipSuccess)) {
    set_error("hipFreeAsync failed: %s",
              hipGetErrorString(fe != hipSuccess ? fe : (fe2 != hipSuccess ? fe2 : fe3)));
    return RTMC_ERR_HIP;
  }
  return rc;
}

// Validates a batch and builds its plan and problem list (items with no
// particles are dropped).  `adj` may be null (forward).
int prepare_batch(const rtmc_batch* bt, const rtmc_adjoint_item* adj, Plan* pl, std::vector<ProbDev>* probs,
                  bool* empty) {
  RTMC_REQUIRE(bt != nullptr, "batch must be given");
  RTMC_REQUIRE(bt->n_items >= 0 && (bt->n_items == 0 || bt->items != nullptr), "items must be given");
  int e = check_shared(bt->geometry, bt->bins, bt->n_times, bt->time_steps, bt->precision);
  if (e) return e;
  *empty = true;
  if (bt->n_items == 0) return RTMC_OK;
  const int ac = bt->items[0].material.acoustics;
  int64_t n_total = 0;
  double ev_weighted = 0.0;
  probs->clear();
  for (int i = 0; i < bt->n_items; ++i) {
    const rtmc_batch_item& it = bt->items[i];
    e = check_item(it.material, it.source, it.n_particles, bt->geometry);
    if (e) {
      std::string msg = rtmc_last_error();
      set_error("item %d: %s", i, msg.c_str());
      return e;
    }
    RTMC_REQUIRE(it.material.acoustics == ac, "item %d: all items must be acoustic or all elastic", i);
    if (adj) {
      const rtmc_adjoint_item& a = adj[i];
      RTMC_REQUIRE(ac, "the adjoint supports acoustic media only");
      RTMC_REQUIRE(a.weight != nullptr && a.grad != nullptr, "item %d: adjoint weight and grad must be given", i);
      RTMC_REQUIRE(a.cv > 0.0 && std::isfinite(a.cv), "item %d: adjoint cv must be > 0", i);
      RTMC_REQUIRE(a.n_dlogsigma >= 2 && a.dlogsigma_dlc != nullptr,
                   "item %d: dlogsigma_dlc needs >= 2 values on [0, pi]", i);
    } else {
      RTMC_REQUIRE(it.E != nullptr || it.n_particles == 0, "item %d: E must be given", i);
    }
    if (it.n_particles == 0 || bt->n_times < 2) continue;
    ProbDev P;
    e = make_prob(it.material, it.source, it.n_particles, it.particle_offset, it.seed, bt->dt, bt->n_times,
                  bt->time_steps, it.E, it.counters, &P);
    if (e) return e;
    if (adj) {
      const rtmc_adjoint_item& a = adj[i];
      P.E = nullptr;
      P.w = a.weight;
      P.grad = a.grad;
      P.dhist = a.dhist;
      P.two_over_c = 2.0 / a.cv;
      P.lam_over_expm1 = P.lamS[0] > 0.0 ? P.lamS[0] / expm1(P.lamS[0]) : 1.0;
      P.dlogS = a.dlogSigma_dlc;
      double Q = it.material.QP;
      P.kq = (P.useQ && std::isfinite(Q)) ? 2.0 * kPi * it.material.frequency * P.dt_sub / (Q * Q) : 0.0;
      P.glc = a.dlogsigma_dlc;
      P.n_glc = a.n_dlogsigma;
    }
    probs->push_back(P);
    n_total += it.n_particles;
    // scattering and removal tests per observation interval: n_sub (lamS + lamK)
    ev_weighted += (double)it.n_particles * (double)P.nsub *
                   (std::max(P.lamS[0], P.lamS[1]) + std::max(P.lamK[0], P.lamK[1]));
  }
  if (probs->empty()) return RTMC_OK;
  // compressed tables in LDS when every problem has them (RTMC_LDS_TABLE=0: never);
  // the forward and its adjoint see the same materials, so they draw the same angles
  int lt_cap = 0;
  if (getenv_int("RTMC_LDS_TABLE", 1) != 0 && bt->bins.direction_variable == RTMC_DIRVAR_DIRZ) {
    bool all = true;
    for (const auto& P : *probs)
      for (int t = 0; t < (ac ? 1 : 4); ++t) {
        all = all && P.M.lm[t] > 0;
        lt_cap = std::max(lt_cap, P.M.lm[t]);
      }
    if (!all) lt_cap = 0;
  }
  e = make_plan(bt->geometry, bt->bins, bt->n_times, bt->dt, bt->time_steps, bt->precision, ac, bt->block_size,
                bt->particles_per_thread, bt->hist_mode, bt->event_batch, n_total, (int)probs->size(), adj != nullptr,
                ev_weighted / (double)n_total, lt_cap, pl);
  if (e) return e;
  if (adj) {
    for (auto& P : *probs) P.dhist_stride = pl->hsize;
  }
  *empty = false;
  return RTMC_OK;
}

}  // namespace
}  // namespace rtmc

using namespace rtmc;

extern "C" int rtmc_forward(const rtmc_problem* prob, uint32_t* E, uint64_t* counters, void* stream) {
  clear_error();
  RTMC_REQUIRE(prob != nullptr && E != nullptr, "problem and E must be given");
  rtmc_batch_item item;
  item.material = prob->material;
  item.source = prob->source;
  item.n_particles = prob->n_particles;
  item.particle_offset = prob->particle_offset;
  item.seed = prob->seed;
  item.E = E;
  item.counters = counters;
  rtmc_batch bt;
  memset(&bt, 0, sizeof(bt));
  bt.geometry = prob->geometry;
  bt.bins = prob->bins;
  bt.n_times = prob->n_times;
  bt.time_steps = prob->time_steps;
  bt.dt = prob->dt;
  bt.precision = prob->precision;
  bt.block_size = prob->block_size;
  bt.particles_per_thread = prob->particles_per_thread;
  bt.hist_mode = prob->hist_mode;
  bt.event_batch = 0;
  bt.n_items = 1;
  bt.items = &item;
  return rtmc_forward_batched(&bt, stream);
}

namespace {

// Device address of `scratch` on each GPU, looked up once per device with that
// device current (host code passes this address, never the array's name).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int noalloc_scratch(void** out) {
  int dev = 0;
  RTMC_HIP(hipGetDevice(&dev));
  RTMC_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d: the no-alloc scratch serves devices 0..%d", dev,
               kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
      return RTMC_ERR_HIP;
    }
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return RTMC_OK;
}

}  // namespace

extern "C" int rtmc_forward_noalloc(const rtmc_problem* prob, uint32_t* E, uint64_t* counters, void* stream) {
  clear_error();
  RTMC_REQUIRE(prob != nullptr && E != nullptr, "problem and E must be given");
  void* base = nullptr;
  int e = noalloc_scratch(&base);
  if (e) return e;
  rtmc_batch_item item;
  item.material = prob->material;
  item.source = prob->source;
  item.n_particles = prob->n_particles;
  item.particle_offset = prob->particle_offset;
  item.seed = prob->seed;
  item.E = E;
  item.counters = counters;
  rtmc_batch bt;
  memset(&bt, 0, sizeof(bt));
  bt.geometry = prob->geometry;
  bt.bins = prob->bins;
  bt.n_times = prob->n_times;
  bt.time_steps = prob->time_steps;
  bt.dt = prob->dt;
  bt.precision = prob->precision;
  bt.block_size = prob->block_size;
  bt.particles_per_thread = prob->particles_per_thread;
  bt.hist_mode = prob->hist_mode;
  bt.n_items = 1;
  bt.items = &item;
  Plan pl;
  std::vector<ProbDev> probs;
  bool empty = true;
  e = prepare_batch(&bt, nullptr, &pl, &probs, &empty);
  if (e || empty) return e;
  Arena ar;
  ar.base = (char*)base;
  ar.cap = (size_t)RTMC_NOALLOC_SCRATCH_BYTES;
  return run_plan(pl, probs, 0, (hipStream_t)stream, &ar);
}

extern "C" int rtmc_forward_batched(const rtmc_batch* batch, void* stream) {
  clear_error();
  Plan pl;
  std::vector<ProbDev> probs;
  bool empty = true;
  int e = prepare_batch(batch, nullptr, &pl, &probs, &empty);
  if (e || empty) return e;
  return run_plan(pl, probs, 0, (hipStream_t)stream);
}

extern "C" int rtmc_adjoint_batched(const rtmc_batch* batch, const rtmc_adjoint_item* adjoint, void* stream) {
  clear_error();
  RTMC_REQUIRE(adjoint != nullptr || (batch && batch->n_items == 0), "adjoint items must be given");
  Plan pl;
  std::vector<ProbDev> probs;
  bool empty = true;
  int e = prepare_batch(batch, adjoint, &pl, &probs, &empty);
  if (e || empty) return e;
  return run_plan(pl, probs, 1, (hipStream_t)stream);
}

extern "C" int rtmc_autotune_batched(const rtmc_batch* batch, double sample_fraction, rtmc_tuning* best,
                                     void* stream) {
  clear_error();
  RTMC_REQUIRE(batch != nullptr && best != nullptr, "batch and best must be given");
  RTMC_REQUIRE(sample_fraction > 0.0 && sample_fraction <= 1.0, "sample_fraction must be in (0, 1]");
  hipStream_t st = (hipStream_t)stream;
  Plan pl0;
  std::vector<ProbDev> probs0;
  bool empty = true;
  int e = prepare_batch(batch, nullptr, &pl0, &probs0, &empty);
  if (e) return e;
  best->block_size = 0;
  best->particles_per_thread = 0;
  best->hist_mode = 0;
  best->event_batch = 0;
  best->ms = 0.0;
  best->ms_default = 0.0;
  if (empty) return RTMC_OK;
  const size_t hbytes = (size_t)pl0.hsize * sizeof(uint32_t) * probs0.size();
  RTMC_REQUIRE(hbytes <= ((size_t)256 << 20), "autotune: %zu bytes of histograms exceed the 256 MiB sample budget",
               hbytes);
  // sample: the first particles of every item, into scratch histograms
  void* scratch = nullptr;
  RTMC_HIP(hipMallocAsync(&scratch, hbytes, st));
  std::vector<rtmc_batch_item> items(batch->items, batch->items + batch->n_items);
  {
    size_t k = 0;
    for (auto& it : items) {
      if (it.n_particles == 0) continue;
      it.n_particles = std::max<int64_t>(std::min<int64_t>(it.n_particles, 4096),
                                         (int64_t)((double)it.n_particles * sample_fraction));
      it.E = (uint32_t*)scratch + (size_t)pl0.hsize * k;
      it.counters = nullptr;
      ++k;
    }
  }
  rtmc_batch sb = *batch;
  sb.items = items.data();

  struct Cand {
    int block, ppt, hist, ev;
  };
  std::vector<Cand> cands;
  const bool lds_possible = pl0.hm == HM_LDS || batch->hist_mode == 2;
  std::vector<int> hists;
  if (batch->hist_mode != 0) hists.push_back(batch->hist_mode);
  else {
    if (lds_possible) hists.push_back(2);
    hists.push_back(1);
    hists.push_back(3);
  }
  for (int h : hists) {
    std::vector<int> blocks = (h == 2) ? std::vector<int>{1024, 512} : std::vector<int>{256, 128};
    for (int b : blocks)
      for (int ev : {8, 16, 24, 32, 40, 48})
        for (int ppt : {0, 4}) cands.push_back({b, ppt, h, ev});
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("autotune: hipEventCreate failed");
    if (e0) (void)hipEventDestroy(e0);
    (void)hipFreeAsync(scratch, st);
    return RTMC_ERR_HIP;
  }
  auto time_one = [&](const Cand* c, double* ms) -> int {
    rtmc_batch tb = sb;
    if (c) {
      tb.block_size = c->block;
      tb.particles_per_thread = c->ppt;
      tb.hist_mode = c->hist;
      tb.event_batch = c->ev;
    }
    float best_ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      RTMC_HIP(hipEventRecord(e0, st));
      int rc = rtmc_forward_batched(&tb, st);
      if (rc) return rc;
      RTMC_HIP(hipEventRecord(e1, st));
      RTMC_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      RTMC_HIP(hipEventElapsedTime(&t, e0, e1));
      best_ms = std::min(best_ms, t);
    }
    *ms = best_ms;
    return RTMC_OK;
  };
  int rc = RTMC_OK;
  double tdef = 0.0;
  rc = time_one(nullptr, &tdef);
  double tbest = tdef;
  // the library default, resolved
  Cand cbest = {(int)pl0.block, batch->particles_per_thread,
                pl0.hm == HM_LDS ? 2 : (pl0.hm == HM_TMAJOR ? 3 : 1), pl0.p.batch};
  for (size_t i = 0; rc == RTMC_OK && i < cands.size(); ++i) {
    double t = 0.0;
    rc = time_one(&cands[i], &t);
    if (rc == RTMC_ERR_ARG) {  // a candidate the problem cannot take (LDS too small): skip it
      clear_error();
      rc = RTMC_OK;
      continue;
    }
    if (rc == RTMC_OK && t < tbest) {
      tbest = t;
      cbest = cands[i];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  hipError_t fe = hipFreeAsync(scratch, st);
  if (rc) return rc;
  if (fe != hipSuccess) {
    set_error("hipFreeAsync failed: %s", hipGetErrorString(fe));
    return RTMC_ERR_HIP;
  }
  best->block_size = cbest.block;
  best->particles_per_thread = cbest.ppt;
  best->hist_mode = cbest.hist;
  best->event_batch = cbest.ev;
  best->ms = tbest;
  best->ms_default = tdef;
  return RTMC_OK;
}

extern "C" int rtmc_set_trace_timing(int32_t on) {
  clear_error();
  g_timer.on = on != 0;
  g_timer.n = 0;
  return RTMC_OK;
}

extern "C" int rtmc_trace_times(double* ms, int32_t max, int32_t* n) {
  clear_error();
  RTMC_REQUIRE(n != nullptr && (ms != nullptr || max == 0), "ms and n must be given");
  TraceTimer& T = g_timer;
  int k = 0;
  for (int i = 0; i < T.n && k < max; ++i, ++k) {
    RTMC_HIP(hipEventSynchronize(T.ev[2 * i + 1]));
    float t = 0.f;
    RTMC_HIP(hipEventElapsedTime(&t, T.ev[2 * i], T.ev[2 * i + 1]));
    ms[k] = t;
  }
  *n = k;
  T.n = 0;
  return RTMC_OK;
}

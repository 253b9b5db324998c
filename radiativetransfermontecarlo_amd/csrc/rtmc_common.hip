// rtmc_common.hip -- error reporting, argument checks and descriptor
// conversion shared by the librtmc entry points.
#include "rtmc_common.hpp"

#include <cstring>
#include <mutex>

namespace rtmc {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

void clear_error() { g_error.clear(); }

const char* last_error() { return g_error.c_str(); }

int get_scratch(Scratch* s) {
  static std::mutex mu;
  static Scratch per_dev[64] = {};
  int dev = 0;
  RTMC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error("device ordinal %d out of range", dev);
    return RTMC_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!per_dev[dev].sum) {
    void* p = nullptr;
    if (hipMalloc(&p, 64) != hipSuccess) {
      set_error("hipMalloc of 64 bytes of scratch failed");
      return RTMC_ERR_NOMEM;
    }
    per_dev[dev].sum = (double*)p;
    per_dev[dev].flag = (int*)((char*)p + 16);
  }
  *s = per_dev[dev];
  return RTMC_OK;
}

int check_geometry(const rtmc_geometry* g) {
  RTMC_REQUIRE(g != nullptr, "geometry is NULL");
  RTMC_REQUIRE(g->dimension == 2 || g->dimension == 3, "geometry.dimension must be 2 or 3 (got %d)", g->dimension);
  RTMC_REQUIRE(g->frame >= 0 && g->frame <= 2, "unknown frame type %d", g->frame);
  RTMC_REQUIRE(g->n_bnd >= 0 && g->n_bnd <= RTMC_MAX_BND, "n_bnd must be in [0, %d]", RTMC_MAX_BND);
  for (int i = 0; i < g->n_bnd; ++i) {
    RTMC_REQUIRE(g->bnd[i].dir >= 1 && g->bnd[i].dir <= 4, "bnd(%d).dir must be 1..4 (got %d)", i + 1,
                 g->bnd[i].dir);
    RTMC_REQUIRE(g->bnd[i].type == RTMC_BND_REFLECTIVE || g->bnd[i].type == RTMC_BND_ABSORBING,
                 "Unknown boundary type %d: use reflective or absorbing", g->bnd[i].type);
  }
  return RTMC_OK;
}

int check_table(const rtmc_table* t, const char* what) {
  if (t->n < 2 || t->x == nullptr || t->y == nullptr) {
    set_error("%s: interpolation table needs >= 2 knots and device pointers", what);
    return RTMC_ERR_TABLE;
  }
  if (t->extrap != RTMC_EXTRAP_LINEAR && t->extrap != RTMC_EXTRAP_NEAREST) {
    set_error("%s: unknown extrapolation rule %d", what, t->extrap);
    return RTMC_ERR_TABLE;
  }
  return RTMC_OK;
}

int check_material(const rtmc_material* m, const rtmc_geometry* g) {
  RTMC_REQUIRE(m != nullptr, "material is NULL");
  int ntab = m->acoustics ? 1 : 4;
  static const char* names[4] = {"invcdf{1,1}", "invcdf{1,2}", "invcdf{2,1}", "invcdf{2,2}"};
  for (int k = 0; k < ntab; ++k) {
    int e = check_table(&m->invcdf[k], names[k]);
    if (e) return e;
  }
  for (int k = 0; k < ntab; ++k) {
    const rtmc_lds_table& L = m->lds[k];
    RTMC_REQUIRE(L.m == 0 || (L.m >= 4 && L.m <= 65535 && L.knots != nullptr),
                 "lds[%d]: a compressed table needs 4..65535 knots and a knot array", k);
    RTMC_REQUIRE(L.m == 0 || L.extrap == m->lds[0].extrap, "lds[%d]: the tables must share their extrapolation", k);
    RTMC_REQUIRE((L.m > 0) == (m->lds[0].m > 0), "lds[%d]: either every table of the material is compressed or none",
                 k);
  }
  RTMC_REQUIRE(m->vp > 0, "material velocity vp must be > 0");
  RTMC_REQUIRE(m->acoustics || m->vs > 0, "material velocity vs must be > 0 for elastic media");
  RTMC_REQUIRE(m->mft[0] > 0 && (m->acoustics || m->mft[1] > 0), "meanFreeTime must be > 0");
  double QS = m->acoustics ? m->QP : m->QS;
  if (std::isfinite(m->QP) || std::isfinite(QS)) {
    RTMC_REQUIRE(!std::isnan(m->frequency),
                 "A finite quality factor Q was provided, but mat.Frequency is empty. "
                 "Please define mat.Frequency before using Q-based attenuation.");
  }
  if (g && !m->acoustics) {
    for (int i = 0; i < g->n_bnd; ++i) {
      if (g->bnd[i].type == RTMC_BND_REFLECTIVE) {
        RTMC_REQUIRE(m->n_zoe >= 2 && m->zoe_deg && m->zoe_Rpp && m->zoe_Rsvsv,
                     "Please define the density in MaterialClass before calling MaterialInterface "
                     "(elastic reflective boundaries need the Zoeppritz tables)");
      }
    }
  }
  return RTMC_OK;
}

MatDev to_matdev(const rtmc_material* m) {
  MatDev M;
  M.acoustics = m->acoustics;
  M.vp = m->vp;
  M.vs = m->acoustics ? 0.0 : m->vs;
  M.mftP = m->mft[0];
  M.mftS = m->acoustics ? m->mft[0] : m->mft[1];
  M.P2P = m->P2P;
  M.S2S = m->S2S;
  for (int k = 0; k < 4; ++k) {
    int src = m->acoustics ? 0 : k;
    M.inv[k].x = (const double*)m->invcdf[src].x;
    M.inv[k].y = (const double*)m->invcdf[src].y;
    M.inv[k].n = m->invcdf[src].n;
    M.inv[k].extrap = m->invcdf[src].extrap;
  }
  M.zdeg = m->zoe_deg;
  M.zRpp = m->zoe_Rpp;
  M.zRss = m->zoe_Rsvsv;
  M.nz = m->n_zoe;
  for (int k = 0; k < 4; ++k) {
    int src = m->acoustics ? 0 : k;
    M.lk[k] = m->lds[src].knots;
    M.lm[k] = m->lds[src].knots ? m->lds[src].m : 0;
  }
  M.lnear = m->lds[0].extrap == RTMC_EXTRAP_NEAREST;
  return M;
}

SlotDraws to_slotdraws(const rtmc_draws* d, int64_t n) {
  SlotDraws D;
  D.kind = d->kind;
  D.n_slots = d->n_slots;
  D.replay = d->replay;
  D.n_steps = d->n_steps;
  D.k0 = (uint32_t)(d->seed & 0xffffffffu);
  D.k1 = (uint32_t)(d->seed >> 32);
  D.stream = (uint32_t)d->stream;
  D.n = n;
  return D;
}

int check_draws(const rtmc_draws* d, int n_slots, const char* what) {
  RTMC_REQUIRE(d != nullptr, "%s: draws is NULL", what);
  RTMC_REQUIRE(d->kind == RTMC_DRAWS_REPLAY || d->kind == RTMC_DRAWS_PHILOX, "%s: unknown draws kind %d", what,
               d->kind);
  if (d->kind == RTMC_DRAWS_REPLAY) {
    if (d->replay == nullptr || d->n_slots != n_slots) {
      set_error("%s: replay draws need %d slots per step (got %d) and a device buffer", what, n_slots, d->n_slots);
      return RTMC_ERR_DRAWS;
    }
  }
  return RTMC_OK;
}

}  // namespace rtmc

extern "C" const char* rtmc_last_error(void) { return rtmc::last_error(); }

extern "C" int rtmc_abi_version(void) { return RTMC_ABI_VERSION; }

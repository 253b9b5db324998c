/*
 * rtmc_oracle.c -- CPU restatement of the reference particle-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of librtmc: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * No product code links, imports or calls it.
 *
 * It restates, in plain C (FP64, libm, no FMA contraction), the MATLAB
 * functions of cottereau/RadiativeTransferMonteCarlo that the GPU kernels
 * replace.  Each function cites the reference lines it follows.  Sub-step
 * loops are written vector-style (outer loop over sub-steps, inner loop over
 * particles) so that the reference's early exits (`if ~any(alive) break`)
 * and the renormalisation of *every* particle at the end of a sub-step are
 * reproduced exactly.
 *
 * Random variates come from the same two sources the GPU kernels accept
 * (include/rtmc.h, rtmc_draws): a replay buffer indexed
 * [step][slot][particle], or Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11), implemented here
 * from the paper independently of the library.
 *
 * The descriptor structs are those of include/rtmc.h; every pointer they
 * hold refers to HOST memory when passed to this file.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/rtmc.h"

#define ORC_PI 3.141592653589793

/* scattering events of the calling thread (read by orc_forward) */
static _Thread_local uint64_t tl_nscat = 0;
/* set by orc_forward when the threads share one histogram (too large to copy per thread) */
static _Thread_local int tl_shared_hist = 0;

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (SC'11 paper, Sec. 2 and Table 2)                     */

static void orc_philox(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0;
        uint32_t n2 = hi0 ^ ctr[3] ^ k1;
        ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

/* exported for the known-answer test */
void orc_philox4x32_10(const uint32_t in_ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {in_ctr[0], in_ctr[1], in_ctr[2], in_ctr[3]};
    orc_philox(c, key[0], key[1]);
    memcpy(out, c, sizeof(c));
}

static double orc_u53(uint32_t w0, uint32_t w1) {
    uint64_t b = ((((uint64_t)w1) << 32) | w0) >> 11;
    return (double)(b + 1) * (1.0 / 9007199254740992.0);
}

/* variate of particle i, draw site `slot`, step `step` (rtmc_draws doc) */
static double orc_draw(const rtmc_draws* D, int64_t n, int64_t i, int slot, int64_t step) {
    if (D->kind == RTMC_DRAWS_REPLAY) {
        return D->replay[(step * D->n_slots + slot) * n + i];
    }
    uint32_t c[4];
    c[0] = (uint32_t)((uint64_t)i & 0xffffffffu);
    c[1] = (uint32_t)(((uint64_t)i >> 32) & 0xffffu) | ((uint32_t)slot << 16);
    c[2] = (uint32_t)step;
    c[3] = (uint32_t)D->stream;
    orc_philox(c, (uint32_t)(D->seed & 0xffffffffu), (uint32_t)(D->seed >> 32));
    return orc_u53(c[0], c[1]);
}

double orc_draw_export(const rtmc_draws* D, int64_t n, int64_t i, int slot, int64_t step) {
    return orc_draw(D, n, i, slot, step);
}

/* ------------------------------------------------------------------ */
/* MATLAB helpers                                                      */

/* griddedInterpolant(x, y) evaluated at q: linear inside, linear or
 * nearest extrapolation outside (MaterialClass.m:2462, initializeParticle.m:201).
 * Interval: largest k <= n-2 with x[k] <= q. */
static double orc_table(const rtmc_table* T, double q) {
    int64_t n = T->n;
    const double* x = T->x;
    const double* y = T->y;
    if (q != q) return q;
    int64_t k;
    if (q < x[0]) {
        if (T->extrap == RTMC_EXTRAP_NEAREST) return y[0];
        k = 0;
    } else if (q > x[n - 1]) {
        if (T->extrap == RTMC_EXTRAP_NEAREST) return y[n - 1];
        k = n - 2;
    } else {
        int64_t lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            int64_t mid = lo + (hi - lo) / 2;
            if (x[mid] <= q) lo = mid; else hi = mid;
        }
        k = lo;
    }
    double t = (q - x[k]) / (x[k + 1] - x[k]);
    return y[k] + t * (y[k + 1] - y[k]);
}

double orc_table_export(const rtmc_table* T, double q) { return orc_table(T, q); }

static double orc_acosd(double c) { return acos(c) * 180.0 / ORC_PI; }

static void orc_cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double orc_sumsq(const double a[3]) { return (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]; }
static double orc_dot(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static double orc_sign(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }

static void ld3(const double* A, int64_t n, int64_t i, double v[3]) { v[0] = A[i]; v[1] = A[n + i]; v[2] = A[2 * n + i]; }
static void st3(double* A, int64_t n, int64_t i, const double v[3]) { A[i] = v[0]; A[n + i] = v[1]; A[2 * n + i] = v[2]; }

/* ------------------------------------------------------------------ */
/* initializeParticle (initializeParticle.m:1-171)                     */

int orc_init_particles(const rtmc_source* S, int32_t acoustics, const rtmc_draws* D,
                       rtmc_particles* P) {
    const int64_t n = P->n;
    const int d = P->dimension;
    for (int64_t i = 0; i < n; ++i) {
        double x[3], dir[3], perp[3];
        if (S->type == RTMC_SRC_POINT) {
            double r;
            if (S->radial.n == 0) {
                double g;
                if (D->kind == RTMC_DRAWS_REPLAY) g = orc_draw(D, n, i, 0, 0);
                else {
                    double u1 = orc_draw(D, n, i, 0, 0), u2 = orc_draw(D, n, i, 8, 0);
                    g = sqrt(-2.0 * log(u1)) * cos(2.0 * ORC_PI * u2);
                }
                r = fabs(g * S->lambda);                                   /* :23 */
            } else {
                r = orc_table(&S->radial, orc_draw(D, n, i, 0, 0));          /* :25-27 */
            }
            double phi0, phi;
            if (d == 2) { phi0 = ORC_PI / 2; phi = ORC_PI / 2; }           /* :31-33 */
            else {
                phi0 = acos(-1.0 + 2.0 * orc_draw(D, n, i, 1, 0));          /* :35 */
                phi = acos(-1.0 + 2.0 * orc_draw(D, n, i, 2, 0));           /* :36 */
            }
            double theta0 = 2.0 * ORC_PI * orc_draw(D, n, i, 3, 0);         /* :38 */
            double theta = 2.0 * ORC_PI * orc_draw(D, n, i, 4, 0);          /* :39 */
            if (S->direction == RTMC_DIR_OUTGOING) { theta = theta0; phi = phi0; }      /* :40-42 */
            else if (S->direction == RTMC_DIR_UPPER) {                                  /* :44-46 */
                theta = 2.0 * ORC_PI * orc_draw(D, n, i, 5, 0);
                phi = acos(orc_draw(D, n, i, 6, 0));
            }
            dir[0] = cos(theta) * sin(phi); dir[1] = sin(theta) * sin(phi); dir[2] = cos(phi);   /* :48 */
            perp[0] = -sin(theta); perp[1] = cos(theta); perp[2] = 0.0;                           /* :49 */
            if (d == 3) {                                                                         /* :51-54 */
                double alpha = 2.0 * ORC_PI * orc_draw(D, n, i, 7, 0);
                double c[3]; orc_cross(dir, perp, c);
                double ca = cos(alpha), sa = sin(alpha);
                for (int k = 0; k < 3; ++k) perp[k] = perp[k] * ca + c[k] * sa;
            }
            double ux = cos(theta0) * sin(phi0), uy = sin(theta0) * sin(phi0), uz = cos(phi0);  /* :57-58 */
            x[0] = S->position[0] + r * ux;
            x[1] = S->position[1] + r * uy;
            x[2] = S->position[2] + r * uz;
        } else {
            /* plane source (:61-147); n, u, v prepared on the host (:67-100) */
            double g;
            if (D->kind == RTMC_DRAWS_REPLAY) g = orc_draw(D, n, i, 0, 0);
            else {
                double u1 = orc_draw(D, n, i, 0, 0), u2 = orc_draw(D, n, i, 8, 0);
                g = sqrt(-2.0 * log(u1)) * cos(2.0 * ORC_PI * u2);
            }
            double c_long = g * S->lambda / 2.0;                                  /* :94 */
            double c1, c2;
            if (S->aperture == RTMC_APERTURE_RECTANGLE) {                         /* :104-119 */
                c1 = (orc_draw(D, n, i, 1, 0) - 0.5) * S->extent[0];
                c2 = (orc_draw(D, n, i, 2, 0) - 0.5) * S->extent[1];
            } else {                                                              /* :121-136 */
                double rho = S->extent[0] * sqrt(orc_draw(D, n, i, 1, 0));
                double alpha = 2.0 * ORC_PI * orc_draw(D, n, i, 2, 0);
                c1 = rho * cos(alpha);
                c2 = rho * sin(alpha);
            }
            for (int k = 0; k < 3; ++k)
                x[k] = ((S->position[k] + c_long * S->normal[k]) + c1 * S->basis_u[k]) + c2 * S->basis_v[k];
            for (int k = 0; k < 3; ++k) dir[k] = S->normal[k];                   /* :143 */
            double th = orc_draw(D, n, i, 4, 0) * 2.0 * ORC_PI;                   /* :146 */
            for (int k = 0; k < 3; ++k) perp[k] = cos(th) * S->basis_u[k] + sin(th) * S->basis_v[k];
        }
        st3(P->x, n, i, x);
        st3(P->dir, n, i, dir);
        st3(P->perp, n, i, perp);
        P->t[i] = 0.0;                                                            /* :152 */
        P->p[i] = (!acoustics && S->polarization == 1) ? 0 : 1;                   /* :157-160 */
        P->alive[i] = 1;                                                          /* :170 */
    }
    return RTMC_OK;
}

/* ------------------------------------------------------------------ */
/* propagateParticleSmallDt (propagateParticleSmallDt.m:1-647)          */

/* MaterialClass.MaterialInterface decision for one particle at a plane
 * boundary of normal e_b (MaterialClass.m:1587-1735).  Returns the branch:
 * 0 P->P, 1 P->SV, 2 SH->SH, 3 SV->SV, 4 SV->P. */
static int orc_interface_plane(const rtmc_material* M, int b, const double dir[3], const double perp[3],
                               int isP, double uA, double uB) {
    double nvec[3] = {0, 0, 0};
    nvec[b] = 1.0;
    rtmc_table zpp = {M->n_zoe, M->zoe_deg, M->zoe_Rpp, RTMC_EXTRAP_LINEAR, 0};
    rtmc_table zss = {M->n_zoe, M->zoe_deg, M->zoe_Rsvsv, RTMC_EXTRAP_LINEAR, 0};
    if (isP) {
        double a = fabs(orc_dot(dir, nvec));                                  /* :1597-1599 */
        if (a > 1) a = 1;
        double inc = orc_acosd(a);                                            /* :1600 */
        double R = orc_table(&zpp, inc);                                      /* :1679 interp1 'extrap' */
        return (uA <= R) ? 0 : 1;                                             /* :1681-1682 */
    }
    double eSH[3];
    orc_cross(dir, nvec, eSH);                                                /* :1613 */
    double nsh = sqrt(orc_sumsq(eSH));
    if (nsh < DBL_EPSILON) {                                                  /* :1616-1625 */
        double t[3] = {1, 0, 0};
        double tn = orc_dot(t, nvec);
        for (int k = 0; k < 3; ++k) t[k] = t[k] - tn * nvec[k];
        for (int k = 0; k < 3; ++k) eSH[k] = t[k];
        nsh = sqrt(orc_sumsq(t));
    }
    for (int k = 0; k < 3; ++k) eSH[k] = eSH[k] / nsh;                        /* :1627 */
    double eSV[3];
    orc_cross(eSH, dir, eSV);                                                 /* :1630 */
    double nsv = sqrt(orc_sumsq(eSV));
    for (int k = 0; k < 3; ++k) eSV[k] = eSV[k] / nsv;                        /* :1631 */
    double aSV = orc_dot(perp, eSV), aSH = orc_dot(perp, eSH);                /* :1634-1635 */
    double ESV = aSV * aSV, ESH = aSH * aSH;
    double Etot = ESV + ESH + DBL_EPSILON;                                    /* :1640 */
    double probSV = ESV / Etot;                                               /* :1643 */
    int isSV = uA < probSV;                                                   /* :1646 */
    if (!isSV) return 2;                                                      /* :1720 SH reflects */
    double a = fabs(orc_dot(dir, nvec));
    if (a > 1) a = 1;
    double inc = orc_acosd(a);                                                /* :1653 */
    double R = orc_table(&zss, inc);                                          /* :1700 */
    return (uB <= R) ? 3 : 4;                                                 /* :1702-1703 */
}

int orc_propagate_smalldt(const rtmc_material* M, const rtmc_geometry* G, rtmc_particles* P,
                          double T, double T_eff_in, const rtmc_draws* D, rtmc_substep_info* info) {
    const int64_t n = P->n;
    const int dim = P->dimension;
    /* :4-7 choice of the sub-step */
    double Te;
    if (T_eff_in != T_eff_in) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += P->t[i];
        Te = T - s / (double)n;
    } else {
        Te = T_eff_in;
    }
    double mftmin = M->acoustics ? M->mft[0] : fmin(M->mft[0], M->mft[1]);
    double dt = fmin(0.01 * mftmin, Te);
    int64_t Nsub;
    if (Te == 0.0) Nsub = 0;
    else Nsub = (int64_t)ceil(Te / dt);
    dt = (Nsub > 0) ? Te / (double)Nsub : 0.0;
    /* :31-82 intrinsic attenuation */
    double QP = M->QP, QS = M->acoustics ? M->QP : M->QS;
    int useQ = isfinite(QP) || isfinite(QS);
    double pKP = 0.0, pKS = 0.0;
    if (useQ) {
        if (!(M->frequency == M->frequency)) return RTMC_ERR_ARG;             /* :59-62 */
        double omega = 2.0 * ORC_PI * M->frequency;
        pKP = isfinite(QP) ? 1.0 - exp(-omega * dt / QP) : 0.0;
        pKS = isfinite(QS) ? 1.0 - exp(-omega * dt / QS) : 0.0;
    }
    if (info) { info->T_eff = Te; info->dt = dt; info->n_sub = Nsub; info->pkill_p = pKP; info->pkill_s = pKS; }
    if (D->kind == RTMC_DRAWS_REPLAY && (D->n_steps < Nsub || D->n_slots != 6 + 2 * G->n_bnd)) return RTMC_ERR_DRAWS;

    double mftP = M->mft[0], mftS = M->acoustics ? M->mft[0] : M->mft[1];       /* :591-593 */
    double* x1 = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    double* theta = (double*)malloc(sizeof(double) * (size_t)n);
    uint8_t* scat = (uint8_t*)malloc((size_t)n);
    uint8_t* wasP = (uint8_t*)malloc((size_t)n);
    if (!x1 || !theta || !scat || !wasP) { free(x1); free(theta); free(scat); free(wasP); return RTMC_ERR_NOMEM; }

    for (int64_t i1 = 0; i1 < Nsub; ++i1) {                                   /* :85 */
        int64_t nalive = 0;
        for (int64_t i = 0; i < n; ++i) nalive += P->alive[i];
        if (nalive == 0) break;                                                /* :91-93 */
        if (useQ) {                                                            /* :100-138 */
            int any = 0;
            for (int64_t i = 0; i < n; ++i) {
                if (!P->alive[i]) continue;
                double pk = (M->acoustics || P->p[i]) ? pKP : pKS;
                if (pk > 0) {
                    double u = orc_draw(D, n, i, 0, i1);
                    if (u < pk) { P->alive[i] = 0; any = 1; }
                }
            }
            if (any) {
                nalive = 0;
                for (int64_t i = 0; i < n; ++i) nalive += P->alive[i];
                if (nalive == 0) break;                                        /* :129-131 */
            }
        }
        /* :141-143 ballistic move */
        memcpy(x1, P->x, sizeof(double) * 3 * (size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            if (!P->alive[i]) continue;
            double v = P->p[i] ? M->vp : M->vs;
            double h = v * dt;
            for (int k = 0; k < 3; ++k) P->x[k * n + i] = P->x[k * n + i] + h * P->dir[k * n + i];
        }
        /* :146-582 boundaries */
        for (int ib = 0; ib < G->n_bnd; ++ib) {
            const rtmc_boundary* B = &G->bnd[ib];
            int absorbing = B->type == RTMC_BND_ABSORBING;
            if (B->dir < 4) {
                int b = B->dir - 1;
                for (int64_t i = 0; i < n; ++i) {
                    if (!P->alive[i]) continue;
                    if (!((P->x[b * n + i] - B->val) * (x1[b * n + i] - B->val) < 0)) continue;   /* :171-172 */
                    if (absorbing) { P->alive[i] = 0; continue; }                                /* :178-182 */
                    if (M->acoustics) {                                                          /* :184-187 */
                        P->x[b * n + i] = (2 * B->val) - P->x[b * n + i];
                        P->dir[b * n + i] = -P->dir[b * n + i];
                        continue;
                    }
                    double dir[3], perp[3];
                    ld3(P->dir, n, i, dir); ld3(P->perp, n, i, perp);
                    double uA = orc_draw(D, n, i, 6 + 2 * ib, i1);
                    double uB = orc_draw(D, n, i, 7 + 2 * ib, i1);
                    int br = orc_interface_plane(M, b, dir, perp, P->p[i], uA, uB);             /* :195 */
                    P->x[b * n + i] = (2 * B->val) - P->x[b * n + i];                           /* :198 */
                    if (br == 0 || br == 2 || br == 3) {                                         /* :204-209, :252-271 */
                        dir[b] = -dir[b]; perp[b] = -perp[b];
                    } else if (br == 1) {                                                        /* :212-248 */
                        P->p[i] = 0;
                        double ratio = M->vs / M->vp;
                        double dn = dir[b];
                        double dtn[3] = {dir[0], dir[1], dir[2]};
                        dtn[b] = 0.0;
                        for (int k = 0; k < 3; ++k) dtn[k] = ratio * dtn[k];
                        double sin2 = orc_sumsq(dtn);
                        double cv = sqrt(fmax(0.0, 1.0 - sin2));
                        dtn[b] = -orc_sign(dn) * cv;
                        double nn = sqrt(orc_sumsq(dtn));
                        for (int k = 0; k < 3; ++k) dir[k] = dtn[k] / nn;
                        double nv[3] = {0, 0, 0}; nv[b] = 1.0;
                        double ax[3]; orc_cross(dir, nv, ax);
                        double na = sqrt(orc_sumsq(ax));
                        for (int k = 0; k < 3; ++k) ax[k] = ax[k] / na;
                        orc_cross(ax, dir, perp);
                    } else {                                                                     /* :274-320 */
                        P->p[i] = 1;
                        double ratio = M->vp / M->vs;
                        double dn = dir[b];
                        double dtn[3] = {dir[0], dir[1], dir[2]};
                        dtn[b] = 0.0;
                        for (int k = 0; k < 3; ++k) dtn[k] = ratio * dtn[k];
                        double sin2 = orc_sumsq(dtn);
                        if (sin2 <= 1) {
                            double cv = sqrt(1 - sin2);
                            dtn[b] = -orc_sign(dn) * cv;
                            double nn = sqrt(orc_sumsq(dtn));
                            for (int k = 0; k < 3; ++k) dir[k] = dtn[k] / nn;
                        } else {
                            P->p[i] = 0;
                            dir[b] = -dir[b];
                        }
                    }
                    st3(P->dir, n, i, dir); st3(P->perp, n, i, perp);
                }
            } else {
                /* :325-580 cylinder of radius val around z */
                rtmc_table zpp = {M->n_zoe, M->zoe_deg, M->zoe_Rpp, RTMC_EXTRAP_NEAREST, 0};
                rtmc_table zss = {M->n_zoe, M->zoe_deg, M->zoe_Rsvsv, RTMC_EXTRAP_NEAREST, 0};
                for (int64_t i = 0; i < n; ++i) {
                    if (!P->alive[i]) continue;
                    double X[3]; ld3(P->x, n, i, X);
                    double r = sqrt(X[0] * X[0] + X[1] * X[1]);                                 /* :328-329 */
                    double ov = r - B->val;                                                      /* :332 */
                    if (!(ov > 0)) continue;                                                     /* :334 */
                    if (absorbing) { P->alive[i] = 0; continue; }
                    double nv[3] = {X[0] / r, X[1] / r, 0.0};                                    /* :348-349 */
                    for (int k = 0; k < 3; ++k) X[k] = X[k] - 2 * ov * nv[k];                    /* :352 */
                    st3(P->x, n, i, X);
                    double d[3], pp[3];
                    ld3(P->dir, n, i, d); ld3(P->perp, n, i, pp);
                    if (M->acoustics) {                                                          /* :354-359 */
                        double dp = orc_dot(d, nv);
                        for (int k = 0; k < 3; ++k) d[k] = d[k] - 2 * dp * nv[k];
                        st3(P->dir, n, i, d);
                        continue;
                    }
                    double uA = orc_draw(D, n, i, 6 + 2 * ib, i1);
                    double uB = orc_draw(D, n, i, 7 + 2 * ib, i1);
                    if (P->p[i]) {                                                               /* :373-455 */
                        double ct = orc_dot(d, nv);
                        double thd = orc_acosd(fmin(1.0, fmax(-1.0, ct)));
                        double R = orc_table(&zpp, thd);
                        if (uA <= R) {
                            double dtn = orc_dot(d, nv);
                            double ptn = orc_dot(pp, nv);
                            for (int k = 0; k < 3; ++k) { d[k] = d[k] - 2 * dtn * nv[k]; pp[k] = pp[k] - 2 * ptn * nv[k]; }
                        } else {
                            P->p[i] = 0;
                            double eta = M->vs / M->vp;
                            double ddn = orc_dot(d, nv);
                            double dt_[3], dnt[3];
                            for (int k = 0; k < 3; ++k) { double dnv = ddn * nv[k]; dt_[k] = d[k] - dnv; }
                            for (int k = 0; k < 3; ++k) dnt[k] = eta * dt_[k];
                            double sin2 = orc_sumsq(dnt);
                            double cv = sqrt(fmax(0.0, 1 - sin2));
                            double dnew[3];
                            for (int k = 0; k < 3; ++k) dnew[k] = dnt[k] - cv * nv[k];
                            double nn = sqrt(orc_sumsq(dnew));
                            for (int k = 0; k < 3; ++k) dnew[k] = dnew[k] / nn;
                            for (int k = 0; k < 3; ++k) d[k] = dnew[k];
                            double ax[3]; orc_cross(dnew, nv, ax);
                            double al = sqrt(orc_sumsq(ax));
                            if (al < 1e-6) { double e1[3] = {1, 0, 0}; orc_cross(dnew, e1, ax); }
                            double an = sqrt(orc_sumsq(ax));
                            for (int k = 0; k < 3; ++k) ax[k] = ax[k] / an;
                            orc_cross(ax, dnew, pp);
                        }
                    } else {                                                                     /* :458-576 */
                        double ai[3]; orc_cross(d, nv, ai);
                        double an = sqrt(orc_sumsq(ai));
                        for (int k = 0; k < 3; ++k) ai[k] = ai[k] / an;
                        double shc = orc_dot(pp, ai);
                        double prob_sh = shc * shc;
                        if (uA < prob_sh) {
                            double dtn = orc_dot(d, nv);
                            double ptn = orc_dot(pp, nv);
                            for (int k = 0; k < 3; ++k) { d[k] = d[k] - 2 * dtn * nv[k]; pp[k] = pp[k] - 2 * ptn * nv[k]; }
                        } else {
                            double ct = orc_dot(d, nv);
                            double thd = orc_acosd(fmin(1.0, fmax(-1.0, ct)));
                            double R = orc_table(&zss, thd);
                            if (uB <= R) {
                                double dtn = orc_dot(d, nv);
                                double ptn = orc_dot(pp, nv);
                                for (int k = 0; k < 3; ++k) { d[k] = d[k] - 2 * dtn * nv[k]; pp[k] = pp[k] - 2 * ptn * nv[k]; }
                            } else {
                                P->p[i] = 1;
                                double eta = M->vp / M->vs;
                                double ddn = orc_dot(d, nv);
                                double dt_[3], dnt[3];
                                for (int k = 0; k < 3; ++k) { double dnv = ddn * nv[k]; dt_[k] = d[k] - dnv; }
                                for (int k = 0; k < 3; ++k) dnt[k] = eta * dt_[k];
                                double sin2 = orc_sumsq(dnt);
                                if (sin2 <= 1) {
                                    double cv = sqrt(1 - sin2);
                                    double dnew[3];
                                    for (int k = 0; k < 3; ++k) dnew[k] = dnt[k] - cv * nv[k];
                                    double nn = sqrt(orc_sumsq(dnew));
                                    for (int k = 0; k < 3; ++k) d[k] = dnew[k] / nn;
                                } else {
                                    P->p[i] = 0;
                                    double dtn = orc_dot(d, nv);
                                    double ptn = orc_dot(pp, nv);
                                    for (int k = 0; k < 3; ++k) { d[k] = d[k] - 2 * dtn * nv[k]; pp[k] = pp[k] - 2 * ptn * nv[k]; }
                                }
                            }
                        }
                    }
                    st3(P->dir, n, i, d); st3(P->perp, n, i, pp);
                }
            }
        }
        /* :585-588 refresh */
        int64_t nal = 0;
        for (int64_t i = 0; i < n; ++i) nal += P->alive[i];
        if (nal == 0) break;
        /* :591-601 Poisson scattering flag */
        for (int64_t i = 0; i < n; ++i) {
            scat[i] = 0;
            wasP[i] = P->p[i];
            if (!P->alive[i]) continue;
            double lam = dt / (P->p[i] ? mftP : mftS);
            double u = orc_draw(D, n, i, 1, i1);
            if (D->kind == RTMC_DRAWS_REPLAY) scat[i] = u > 0.0;
            else scat[i] = u < 1.0 - exp(-lam);
            tl_nscat += scat[i];
        }
        /* :604-611 azimuth */
        for (int64_t i = 0; i < n; ++i) {
            if (!scat[i]) continue;
            double pp[3]; ld3(P->perp, n, i, pp);
            if (dim == 2) {
                double u = orc_draw(D, n, i, 2, i1);
                int bit = (D->kind == RTMC_DRAWS_REPLAY) ? (u != 0.0) : (u < 0.5);
                double s = bit ? -1.0 : 1.0;
                for (int k = 0; k < 3; ++k) pp[k] = s * pp[k];
            } else {
                double phi = (2 * ORC_PI) * orc_draw(D, n, i, 2, i1);
                double d[3], c[3];
                ld3(P->dir, n, i, d);
                orc_cross(d, pp, c);
                double cp = cos(phi), sp = sin(phi);
                for (int k = 0; k < 3; ++k) pp[k] = cp * pp[k] + sp * c[k];
            }
            st3(P->perp, n, i, pp);
        }
        /* :614-632 polar angle and polarisation */
        for (int64_t i = 0; i < n; ++i) {
            theta[i] = 0.0;
            if (!scat[i]) continue;
            double u = orc_draw(D, n, i, 3, i1);
            if (M->acoustics) {
                theta[i] = orc_table(&M->invcdf[0], u);
            } else {
                int from_p = wasP[i];
                if (from_p) { if (orc_draw(D, n, i, 4, i1) > M->P2P) P->p[i] = 0; }
                else { if (orc_draw(D, n, i, 5, i1) > M->S2S) P->p[i] = 1; }
                int tab = from_p ? (P->p[i] ? 0 : 1) : (P->p[i] ? 2 : 3);
                theta[i] = orc_table(&M->invcdf[tab], u);
            }
        }
        /* :635-639 rotation and renormalisation of all particles */
        for (int64_t i = 0; i < n; ++i) {
            double d[3], pp[3], nd[3], np[3];
            ld3(P->dir, n, i, d); ld3(P->perp, n, i, pp);
            double ct = cos(theta[i]), st = sin(theta[i]);
            for (int k = 0; k < 3; ++k) nd[k] = ct * d[k] + st * pp[k];
            double nrm = sqrt(orc_sumsq(nd));
            for (int k = 0; k < 3; ++k) nd[k] = nd[k] / nrm;
            for (int k = 0; k < 3; ++k) np[k] = -st * d[k] + ct * pp[k];
            double nrp = sqrt(orc_sumsq(np));
            for (int k = 0; k < 3; ++k) np[k] = np[k] / nrp;
            st3(P->dir, n, i, nd); st3(P->perp, n, i, np);
        }
    }
    for (int64_t i = 0; i < n; ++i) P->t[i] = P->t[i] + Te;                     /* :645 */
    free(x1); free(theta); free(scat); free(wasP);
    return RTMC_OK;
}

/* ------------------------------------------------------------------ */
/* propagateParticle (propagateParticle.m:1-96)                         */

static int64_t orc_poisson(double lam, double u) {
    /* inversion of the Poisson CDF with one uniform (PHILOX draws) */
    int64_t k = 0;
    double pk = exp(-lam), F = pk;
    while (u > F && pk > 0.0) {
        k += 1;
        pk = pk * lam / (double)k;
        F += pk;
    }
    return k;
}

int orc_propagate_events(const rtmc_material* M, rtmc_particles* P, double T, const rtmc_draws* D) {
    const int64_t n = P->n;
    const int dim = P->dimension;
    if (D->kind == RTMC_DRAWS_REPLAY && D->n_slots != 6) return RTMC_ERR_DRAWS;
    for (int64_t i = 0; i < n; ++i) {
        double dt = T - P->t[i];                                              /* :4 */
        int64_t it = 0;
        while (dt > 0) {                                                      /* :8 */
            if (D->kind == RTMC_DRAWS_REPLAY && it >= D->n_steps) return RTMC_ERR_DRAWS;
            int isP = P->p[i];
            double mft = isP ? M->mft[0] : (M->acoustics ? M->mft[0] : M->mft[1]);
            int64_t Nj;                                                       /* :15-19 */
            double u0 = orc_draw(D, n, i, 0, it);
            if (D->kind == RTMC_DRAWS_REPLAY) Nj = (int64_t)u0;
            else {
                double lam = dt / mft;
                if (lam > 700.0) return RTMC_ERR_ARG;
                Nj = orc_poisson(lam, u0);
            }
            int sc = Nj > 0;                                                  /* :20 */
            tl_nscat += (uint64_t)sc;
            if (sc) {                                                         /* :24-26, :86-96 */
                double umin = orc_draw(D, n, i, 1, it);
                if (D->kind != RTMC_DRAWS_REPLAY) umin = 1.0 - pow(umin, 1.0 / (double)Nj);
                dt = umin * dt;
            }
            double v = isP ? M->vp : M->vs;                                   /* :29-30 */
            double h = v * dt;
            for (int k = 0; k < 3; ++k) P->x[k * n + i] = P->x[k * n + i] + h * P->dir[k * n + i];
            if (sc) {
                double d[3], pp[3];
                ld3(P->dir, n, i, d); ld3(P->perp, n, i, pp);
                if (dim == 2) {                                               /* :33-35 */
                    double u = orc_draw(D, n, i, 2, it);
                    int bit = (D->kind == RTMC_DRAWS_REPLAY) ? (u != 0.0) : (u < 0.5);
                    double s = bit ? -1.0 : 1.0;
                    for (int k = 0; k < 3; ++k) pp[k] = s * pp[k];
                } else {                                                      /* :36-39 */
                    double phi = (2 * ORC_PI) * orc_draw(D, n, i, 2, it);
                    double c[3]; orc_cross(d, pp, c);
                    double cp = cos(phi), sp = sin(phi);
                    for (int k = 0; k < 3; ++k) pp[k] = cp * pp[k] + sp * c[k];
                }
                double th;
                double u3 = orc_draw(D, n, i, 3, it);
                if (M->acoustics) th = orc_table(&M->invcdf[0], u3);          /* :46 */
                else {                                                        /* :51-63 */
                    if (isP) { if (orc_draw(D, n, i, 4, it) > M->P2P) P->p[i] = 0; }
                    else { if (orc_draw(D, n, i, 5, it) > M->S2S) P->p[i] = 1; }
                    int tab = isP ? (P->p[i] ? 0 : 1) : (P->p[i] ? 2 : 3);
                    th = orc_table(&M->invcdf[tab], u3);
                }
                double ct = cos(th), st = sin(th), nd[3], np[3];              /* :68-72 */
                for (int k = 0; k < 3; ++k) nd[k] = ct * d[k] + st * pp[k];
                double nrm = sqrt(orc_sumsq(nd));
                for (int k = 0; k < 3; ++k) nd[k] = nd[k] / nrm;
                for (int k = 0; k < 3; ++k) np[k] = -st * d[k] + ct * pp[k];
                double nrp = sqrt(orc_sumsq(np));
                for (int k = 0; k < 3; ++k) np[k] = np[k] / nrp;
                st3(P->dir, n, i, nd); st3(P->perp, n, i, np);
            }
            P->t[i] = P->t[i] + dt;                                           /* :75 */
            dt = T - P->t[i];                                                 /* :76 */
            it += 1;
        }
    }
    return RTMC_OK;
}

/* ------------------------------------------------------------------ */
/* observeTime (observeTime.m:1-146) + E(:,:,it,:) += ...               */

/* histcounts bin of v over edges e[0..ne-1]: e_k <= v < e_k+1, the last
 * bin closed; -1 outside or NaN. */
static int64_t orc_bin(const double* e, int ne, double v) {
    if (!(v >= e[0] && v <= e[ne - 1])) return -1;
    int64_t lo = 0, hi = ne - 1;
    while (hi - lo > 1) {
        int64_t mid = lo + (hi - lo) / 2;
        if (e[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

int64_t orc_bin_export(const double* e, int ne, double v) { return orc_bin(e, ne, v); }

/* the four observation variables of one particle (observeTime.m:15-45) */
static void orc_obs_vars(int dim, int frame, const double X[3], double dz, double vars[4]) {
    double r, th, ph = 0.0;
    if (dim == 2) {
        th = atan2(X[1], X[0]);                                               /* cart2pol :15 */
        r = hypot(X[0], X[1]);
    } else {
        double hxy = hypot(X[0], X[1]);                                       /* cart2sph :18 */
        r = hypot(hxy, X[2]);
        ph = atan2(X[2], hxy);
        th = atan2(X[1], X[0]);
    }
    double cp = dz;                                                           /* :24-29 */
    if (cp > 1) cp = 1;
    if (cp < -1) cp = -1;
    double psi = (dim == 3) ? cp : acos(cp);                                  /* :30-35 */
    if (dim == 3 && frame == RTMC_FRAME_CYLINDRICAL) {                        /* :38-40 */
        th = atan2(X[1], X[0]);
        r = hypot(X[0], X[1]);
    }
    if (dim == 3 && frame == RTMC_FRAME_SPHERICAL) ph = sin(ph);              /* :43-45 */
    if (frame == RTMC_FRAME_CARTESIAN) { vars[0] = X[0]; vars[1] = X[1]; vars[2] = X[2]; }
    else if (frame == RTMC_FRAME_CYLINDRICAL) { vars[0] = r; vars[1] = th; vars[2] = X[2]; }
    else { vars[0] = r; vars[1] = th; vars[2] = ph; }
    vars[3] = psi;
}

int orc_observe(const rtmc_geometry* G, const rtmc_bins* B, int32_t acoustics,
                const rtmc_particles* P, uint32_t* E, int64_t time_index) {
    const int64_t n = P->n;
    const int dim = P->dimension;
    for (int64_t i = 0; i < n; ++i) {
        if (!P->alive[i]) continue;                                           /* :138 */
        double X[3]; ld3(P->x, n, i, X);
        double vars[4];
        double dz = P->dir[2 * n + i];
        if (B->direction_variable == RTMC_DIRVAR_POSITION) {
            /* the angle to the position vector (README.md:209), as its cosine */
            double D[3]; ld3(P->dir, n, i, D);
            double dot = D[0] * X[0] + D[1] * X[1] + D[2] * X[2];
            dz = dot / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        }
        orc_obs_vars(dim, G->frame, X, dz, vars);
        int64_t idx = time_index * B->time_stride;
        int ok = 1;
        for (int v = 0; v < 4 && ok; ++v) {
            if (dim == 2 && v == 2) continue;
            int64_t k = orc_bin(B->edges[v], B->n_edges[v], vars[v]);
            if (k < 0) ok = 0; else idx += k * B->stride[v];
        }
        if (!ok) continue;
        if (!acoustics && !P->p[i]) idx += B->mode_stride;                    /* :144-146 */
        if (tl_shared_hist) __atomic_fetch_add(&E[idx], 1u, __ATOMIC_RELAXED);
        else E[idx] += 1u;
    }
    return RTMC_OK;
}

/* ------------------------------------------------------------------ */
/* Whole solve: the packet x time loop of radiativeTransfer             */
/* (radiativeTransfer.m:95-158, serial branch), sub-stepped propagation */
/* (timeSteps = 0) or event-driven (timeSteps = 1), PHILOX draws.       */
/* Packets run on OpenMP threads; each packet draws from its own Philox */
/* key (seed + packet), so the result does not depend on the threads.   */

int orc_forward(const rtmc_problem* prob, int64_t packet, uint32_t* E, int32_t n_threads,
                uint64_t* counters) {
    const int64_t N = prob->n_particles;
    if (packet <= 0) packet = 100000;
    const int64_t Np = (N + packet - 1) / packet;
    const int dim = prob->geometry.dimension;
    const int nmode = prob->material.acoustics ? 1 : 2;
    int64_t hsize = prob->bins.time_stride * prob->n_times;
    if (!prob->material.acoustics) {
        int64_t m2 = prob->bins.mode_stride * 2;
        if (m2 > hsize) hsize = m2;
    }
    int err = RTMC_OK;
    uint64_t tot_alive_deposits = 0, tot_scat = 0;
    (void)nmode;
    /* thread-private histograms (summed at the end, like parfor's reduction of
       E_local, radiativeTransfer.m:65-88) unless they would take over 256 MiB
       each: then every thread adds into E with atomic increments */
    const int shared = (size_t)hsize * sizeof(uint32_t) > ((size_t)256 << 20);
    uint64_t dep_shared0 = 0;
    if (shared)
        for (int64_t k = 0; k < hsize; ++k) dep_shared0 += E[k];
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        tl_shared_hist = shared;
        uint32_t* El = shared ? E : (uint32_t*)calloc((size_t)hsize, sizeof(uint32_t));
        double* buf = (double*)malloc(sizeof(double) * (size_t)(10 * packet + 16));
        uint8_t* bb = (uint8_t*)malloc((size_t)(2 * packet + 16));
        uint64_t dep = 0;
        tl_nscat = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t ip = 0; ip < Np; ++ip) {
            int64_t n = packet;
            if (ip == Np - 1) n = N - ip * packet;
            rtmc_particles P;
            P.n = n; P.dimension = dim;
            P.x = buf; P.dir = buf + 3 * n; P.perp = buf + 6 * n; P.t = buf + 9 * n;
            P.p = bb; P.alive = bb + n;
            uint64_t key = prob->seed + 0x9E3779B97F4A7C15ull * (uint64_t)(prob->particle_offset / (uint64_t)packet + (uint64_t)ip + 1);
            rtmc_draws D = {RTMC_DRAWS_PHILOX, 0, NULL, 0, key, 0};
            orc_init_particles(&prob->source, prob->material.acoustics, &D, &P);
            for (int it = 1; it < prob->n_times; ++it) {                      /* :114 */
                double tval = 0.0 + (double)it * prob->dt;                     /* :117 */
                D.stream = (uint64_t)it;
                int e;
                if (prob->time_steps == 0)
                    e = orc_propagate_smalldt(&prob->material, &prob->geometry, &P, tval, NAN, &D, NULL);
                else
                    e = orc_propagate_events(&prob->material, &P, tval, &D);
                if (e != RTMC_OK) { err = e; break; }
                orc_observe(&prob->geometry, &prob->bins, prob->material.acoustics, &P, El, it);
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (!shared) {
                for (int64_t k = 0; k < hsize; ++k) { E[k] += El[k]; dep += El[k]; }
                free(El);
            }
            tot_alive_deposits += dep;
            tot_scat += tl_nscat;
        }
        tl_shared_hist = 0;
        free(buf); free(bb);
    }
    if (shared) {
        uint64_t s = 0;
        for (int64_t k = 0; k < hsize; ++k) s += E[k];
        tot_alive_deposits = s - dep_shared0;
    }
    if (counters) {
        counters[RTMC_CNT_SCATTER] = tot_scat;
        counters[RTMC_CNT_DEPOSIT] = tot_alive_deposits;
    }
    return err;
}

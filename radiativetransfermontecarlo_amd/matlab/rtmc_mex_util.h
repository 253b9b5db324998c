/*
 * rtmc_mex_util.h -- helpers shared by the MATLAB gateways of librtmc
 * (rtmc_mex.c, rtmc_batch_mex.c): reading the reference's MATLAB structs and
 * MaterialClass objects into the rtmc_* descriptors of include/rtmc.h, and
 * device uploads that are released before any error returns to MATLAB.
 *
 * Built by MATLAB's mex, not by this repository's make (mex.h comes with
 * MATLAB).  Every mx* call happens on MATLAB's thread.
 */
#ifndef RTMC_MEX_UTIL_H
#define RTMC_MEX_UTIL_H

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "mex.h"
#include "rtmc.h"

/* Device allocations of one gateway call (and library handles of material
 * tables, rtmc_free_material); freed on every exit path. */
#define RTMC_MEX_MAX_KEEP 256
#define RTMC_MEX_MAX_HANDLES 8
typedef struct {
    void* p[RTMC_MEX_MAX_KEEP];
    int n;
    void* h[RTMC_MEX_MAX_HANDLES];
    int nh;
} rtmc_mex_keep;

static void rtmc_mex_release(rtmc_mex_keep* k) {
    for (int i = 0; i < k->n; ++i) (void)hipFree(k->p[i]);
    k->n = 0;
    for (int i = 0; i < k->nh; ++i) (void)rtmc_free_material(k->h[i]);
    k->nh = 0;
}

/* Frees the call's device memory, then raises the MATLAB error (does not return). */
static void rtmc_mex_fail(rtmc_mex_keep* k, const char* id, const char* fmt, ...) {
    char msg[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    if (k) rtmc_mex_release(k);
    mexErrMsgIdAndTxt(id, "%s", msg);
}

static void rtmc_mex_check(rtmc_mex_keep* k, int status, const char* what) {
    if (status != RTMC_OK) rtmc_mex_fail(k, "rtmc:call", "%s failed (%d): %s", what, status, rtmc_last_error());
}

static void* rtmc_mex_device(rtmc_mex_keep* k, size_t bytes) {
    void* d = NULL;
    if (k->n >= RTMC_MEX_MAX_KEEP) rtmc_mex_fail(k, "rtmc:nomem", "too many device arrays in one call");
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&d, bytes) != hipSuccess) rtmc_mex_fail(k, "rtmc:nomem", "hipMalloc of %zu bytes failed", bytes);
    k->p[k->n++] = d;
    return d;
}

static const double* rtmc_mex_upload(rtmc_mex_keep* k, const double* host, size_t n) {
    void* d = rtmc_mex_device(k, n * sizeof(double));
    if (n && hipMemcpy(d, host, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rtmc_mex_fail(k, "rtmc:hip", "hipMemcpy to the device failed");
    return (const double*)d;
}

/* Field of a struct, or property of an object (NULL if absent or empty). */
static const mxArray* rtmc_mex_get(const mxArray* a, mwIndex i, const char* name) {
    const mxArray* v = NULL;
    if (a == NULL) return NULL;
    if (mxIsStruct(a)) {
        if (mxGetFieldNumber(a, name) < 0) return NULL;
        v = mxGetField(a, i, name);
    } else {
        v = mxGetProperty(a, i, name);
    }
    if (v == NULL || mxIsEmpty(v)) return NULL;
    return v;
}

static double rtmc_mex_num(rtmc_mex_keep* k, const mxArray* v, size_t idx, const char* what) {
    if (v == NULL || (!mxIsNumeric(v) && !mxIsLogical(v))) rtmc_mex_fail(k, "rtmc:arg", "%s must be numeric", what);
    if (idx >= mxGetNumberOfElements(v)) rtmc_mex_fail(k, "rtmc:arg", "%s has fewer than %zu values", what, idx + 1);
    if (mxIsDouble(v)) return mxGetDoubles(v)[idx];
    {
        mxArray* in[1];
        mxArray* out[1];
        double r;
        in[0] = (mxArray*)v;
        mexCallMATLAB(1, out, 1, in, "double");
        r = mxGetDoubles(out[0])[idx];
        mxDestroyArray(out[0]);
        return r;
    }
}

/* Scalar field/property with a default when absent. */
static double rtmc_mex_scalar(rtmc_mex_keep* k, const mxArray* a, const char* name, double dflt) {
    const mxArray* v = rtmc_mex_get(a, 0, name);
    return v ? rtmc_mex_num(k, v, 0, name) : dflt;
}

/* Lower-cased char field/property into buf; "" when absent. */
static const char* rtmc_mex_string(rtmc_mex_keep* k, const mxArray* a, const char* name, char* buf, size_t len) {
    const mxArray* v = rtmc_mex_get(a, 0, name);
    buf[0] = 0;
    if (v == NULL) return buf;
    if (!mxIsChar(v)) rtmc_mex_fail(k, "rtmc:arg", "%s must be a char array", name);
    if (mxGetString(v, buf, len) != 0) rtmc_mex_fail(k, "rtmc:arg", "%s is too long", name);
    for (char* c = buf; *c; ++c)
        if (*c >= 'A' && *c <= 'Z') *c = (char)(*c - 'A' + 'a');
    return buf;
}

/* Double vector of a MATLAB array (any numeric class), copied into out[0..n). */
static size_t rtmc_mex_vector(rtmc_mex_keep* k, const mxArray* v, double* out, size_t max, const char* what) {
    size_t n;
    if (v == NULL) rtmc_mex_fail(k, "rtmc:arg", "%s is missing", what);
    n = mxGetNumberOfElements(v);
    if (n > max) rtmc_mex_fail(k, "rtmc:arg", "%s has more than %zu values", what, max);
    for (size_t i = 0; i < n; ++i) out[i] = rtmc_mex_num(k, v, i, what);
    return n;
}

/* geometry struct (radiativeTransfer.m:5-9, propagateParticleSmallDt.m:152-165) */
static void rtmc_mex_geometry(rtmc_mex_keep* k, const mxArray* g, rtmc_geometry* out) {
    char frame[32];
    const mxArray* bnd;
    memset(out, 0, sizeof(*out));
    out->dimension = (int32_t)rtmc_mex_scalar(k, g, "dimension", 0);
    if (out->dimension != 2 && out->dimension != 3) rtmc_mex_fail(k, "rtmc:arg", "geometry.dimension must be 2 or 3");
    rtmc_mex_string(k, g, "frame", frame, sizeof(frame));
    if (frame[0] == 0 || strcmp(frame, "spherical") == 0) out->frame = RTMC_FRAME_SPHERICAL;  /* :7-9 */
    else if (strcmp(frame, "cartesian") == 0) out->frame = RTMC_FRAME_CARTESIAN;
    else if (strcmp(frame, "cylindrical") == 0) out->frame = RTMC_FRAME_CYLINDRICAL;
    else rtmc_mex_fail(k, "rtmc:arg", "unknown frame type");
    bnd = rtmc_mex_get(g, 0, "bnd");
    if (bnd != NULL) {
        size_t nb = mxGetNumberOfElements(bnd);
        if (!mxIsStruct(bnd)) rtmc_mex_fail(k, "rtmc:arg", "geometry.bnd must be a struct array");
        if (nb > RTMC_MAX_BND) rtmc_mex_fail(k, "rtmc:arg", "at most %d boundaries are supported", RTMC_MAX_BND);
        out->n_bnd = (int32_t)nb;
        for (size_t i = 0; i < nb; ++i) {
            const mxArray* t = mxGetFieldNumber(bnd, "type") >= 0 ? mxGetField(bnd, i, "type") : NULL;
            char typ[32] = "reflective";
            if (t != NULL && !mxIsEmpty(t)) {
                if (!mxIsChar(t) || mxGetString(t, typ, sizeof(typ)) != 0)
                    rtmc_mex_fail(k, "rtmc:arg", "geometry.bnd(%zu).type must be a char array", i + 1);
                for (char* c = typ; *c; ++c)
                    if (*c >= 'A' && *c <= 'Z') *c = (char)(*c - 'A' + 'a');
            }
            if (strcmp(typ, "reflective") == 0) out->bnd[i].type = RTMC_BND_REFLECTIVE;
            else if (strcmp(typ, "absorbing") == 0) out->bnd[i].type = RTMC_BND_ABSORBING;
            else rtmc_mex_fail(k, "rtmc:arg", "Unknown boundary type \"%s\". Use \"reflective\" or \"absorbing\".", typ);
            out->bnd[i].dir = (int32_t)rtmc_mex_num(k, mxGetField(bnd, i, "dir"), 0, "geometry.bnd.dir");
            out->bnd[i].val = rtmc_mex_num(k, mxGetField(bnd, i, "val"), 0, "geometry.bnd.val");
        }
    }
}

/* The local inverseCDF of initializeParticle.m:175-203 for a radial law given
 * as a griddedInterpolant f: Nth = 10000 radii on [0, Rmax], pdf r^(d-1) f(r),
 * cumulative sum, strictly increasing knots; uploaded as a nearest-extrapolating
 * table. */
static void rtmc_mex_radial(rtmc_mex_keep* k, const mxArray* f, int d, rtmc_table* out) {
    enum { NTH = 10000 };
    static double xth[NTH], fx[NTH], cdf[NTH], ux[NTH], uc[NTH];
    const mxArray* gv = rtmc_mex_get(f, 0, "GridVectors");
    double rmax, sum = 0.0, integ = 0.0, step;
    mxArray* in[2];
    mxArray* o[1];
    size_t m = 0;
    if (gv == NULL || !mxIsCell(gv)) rtmc_mex_fail(k, "rtmc:arg", "source.radial must be a griddedInterpolant");
    {
        const mxArray* g0 = mxGetCell(gv, 0);
        rmax = rtmc_mex_num(k, g0, mxGetNumberOfElements(g0) - 1, "source.radial grid");
    }
    for (int i = 0; i < NTH; ++i) xth[i] = (i == NTH - 1) ? rmax : rmax * (double)i / (double)(NTH - 1);
    in[0] = (mxArray*)f;
    in[1] = mxCreateDoubleMatrix(1, NTH, mxREAL);
    memcpy(mxGetDoubles(in[1]), xth, sizeof(xth));
    mexCallMATLAB(1, o, 2, in, "feval");                       /* f(xth) */
    for (int i = 0; i < NTH; ++i) fx[i] = rtmc_mex_num(k, o[0], (size_t)i, "source.radial(r)") * pow(xth[i], d - 1);
    mxDestroyArray(o[0]);
    mxDestroyArray(in[1]);
    step = 0.0;
    for (int i = 1; i < NTH; ++i) {
        integ += (xth[i] - xth[i - 1]) * (fx[i - 1] + fx[i]) / 2.0;   /* trapz(xth, g) */
        step += xth[i] - xth[i - 1];
    }
    step /= (double)(NTH - 1);                                        /* mean(diff(xth)) */
    for (int i = 0; i < NTH; ++i) {
        sum += fx[i] / integ;
        cdf[i] = sum * step;
    }
    cdf[0] = 0.0;
    cdf[NTH - 1] = 1.0;
    /* knots i and i+1 of every strict increase; cdf is non-decreasing, so its sort is
       itself and the unique values keep their first radius */
    for (int i = 0; i < NTH; ++i) {
        const int keep = (i + 1 < NTH && cdf[i + 1] > cdf[i]) || (i > 0 && cdf[i] > cdf[i - 1]);
        if (!keep) continue;
        if (m > 0 && cdf[i] == uc[m - 1]) continue;
        uc[m] = cdf[i];
        ux[m] = xth[i];
        ++m;
    }
    if (m < 2) rtmc_mex_fail(k, "rtmc:arg", "source.radial gives fewer than two distinct cdf values");
    out->n = (int64_t)m;
    out->x = rtmc_mex_upload(k, uc, m);
    out->y = rtmc_mex_upload(k, ux, m);
    out->extrap = RTMC_EXTRAP_NEAREST;
}

/* source struct (initializeParticle.m:3-160) */
static void rtmc_mex_source(rtmc_mex_keep* k, const mxArray* s, int d, rtmc_source* out) {
    char typ[32], dirn[32], pol[8], ap[32];
    double pos[3] = {0, 0, 0};
    const mxArray* pv;
    memset(out, 0, sizeof(*out));
    pv = rtmc_mex_get(s, 0, "position");
    if (pv != NULL) {
        size_t np = rtmc_mex_vector(k, pv, pos, 3, "source.position");
        if (d == 2 && np == 2) pos[2] = 0.0;
    }
    memcpy(out->position, pos, sizeof(pos));
    out->lambda = rtmc_mex_scalar(k, s, "lambda", 0.0);
    rtmc_mex_string(k, s, "polarization", pol, sizeof(pol));
    out->polarization = strcmp(pol, "s") == 0;
    rtmc_mex_string(k, s, "type", typ, sizeof(typ));
    if (typ[0] == 0 || strcmp(typ, "point") == 0) {
        const mxArray* dv = rtmc_mex_get(s, 0, "direction");
        const mxArray* radial = rtmc_mex_get(s, 0, "radial");
        out->type = RTMC_SRC_POINT;
        out->direction = RTMC_DIR_UNIFORM;
        if (dv != NULL && mxIsChar(dv)) {
            rtmc_mex_string(k, s, "direction", dirn, sizeof(dirn));
            if (strcmp(dirn, "outgoing") == 0) out->direction = RTMC_DIR_OUTGOING;
            else if (strcmp(dirn, "upper") == 0) out->direction = RTMC_DIR_UPPER;
        }
        if (radial != NULL) rtmc_mex_radial(k, radial, d, &out->radial);
    } else if (strcmp(typ, "plane") == 0) {
        double n[3], nn;
        const mxArray* dv = rtmc_mex_get(s, 0, "direction");
        const mxArray* ev = rtmc_mex_get(s, 0, "extent");
        mxArray* in[1];
        mxArray* o[1];
        if (d != 3) rtmc_mex_fail(k, "rtmc:arg", "plane sources are only available in 3D");
        out->type = RTMC_SRC_PLANE;
        rtmc_mex_string(k, s, "aperture", ap, sizeof(ap));
        if (ap[0] == 0 || strcmp(ap, "rectangle") == 0) out->aperture = RTMC_APERTURE_RECTANGLE;
        else if (strcmp(ap, "circle") == 0) out->aperture = RTMC_APERTURE_CIRCLE;
        else rtmc_mex_fail(k, "rtmc:arg", "Unknown plane-wave aperture: use 'rectangle' or 'circle'.");
        if (dv == NULL) rtmc_mex_fail(k, "rtmc:arg", "source.direction is required for a plane source");
        if (mxGetNumberOfElements(dv) == 1) {
            const double a = rtmc_mex_num(k, dv, 0, "source.direction");
            const int ia = (int)a;
            if ((double)ia != a || ia == 0 || ia < -3 || ia > 3)
                rtmc_mex_fail(k, "rtmc:arg", "For a scalar plane-source direction, use +-1, +-2, or +-3.");
            n[0] = n[1] = n[2] = 0.0;
            n[(ia > 0 ? ia : -ia) - 1] = ia > 0 ? 1.0 : -1.0;
        } else if (mxGetNumberOfElements(dv) == 3) {
            rtmc_mex_vector(k, dv, n, 3, "source.direction");
        } else {
            rtmc_mex_fail(k, "rtmc:arg", "For a plane source, source.direction must be a 3-component vector or a "
                                         "scalar axis +-1, +-2, +-3.");
        }
        nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int i = 0; i < 3; ++i) out->normal[i] = n[i] / nn;
        /* null(n) as MATLAB computes it (initializeParticle.m:98-100): the basis that
           orients a rectangular aperture */
        in[0] = mxCreateDoubleMatrix(1, 3, mxREAL);
        memcpy(mxGetDoubles(in[0]), out->normal, 3 * sizeof(double));
        mexCallMATLAB(1, o, 1, in, "null");
        if (mxGetM(o[0]) != 3 || mxGetN(o[0]) != 2) rtmc_mex_fail(k, "rtmc:arg", "null(n) is not 3 x 2");
        for (int i = 0; i < 3; ++i) {
            out->basis_u[i] = mxGetDoubles(o[0])[i];
            out->basis_v[i] = mxGetDoubles(o[0])[3 + i];
        }
        mxDestroyArray(o[0]);
        mxDestroyArray(in[0]);
        if (ev == NULL) rtmc_mex_fail(k, "rtmc:arg", "source.extent is required for a plane source");
        if (out->aperture == RTMC_APERTURE_RECTANGLE) {
            out->extent[0] = rtmc_mex_num(k, ev, 0, "source.extent");
            out->extent[1] = mxGetNumberOfElements(ev) == 1 ? out->extent[0] : rtmc_mex_num(k, ev, 1, "source.extent");
        } else {
            if (mxGetNumberOfElements(ev) != 1)
                rtmc_mex_fail(k, "rtmc:arg", "For source.aperture = 'circle', source.extent must be a scalar radius.");
            out->extent[0] = rtmc_mex_num(k, ev, 0, "source.extent");
        }
    } else {
        rtmc_mex_fail(k, "rtmc:arg", "unknown source type '%s'", typ);
    }
}

/* obs and ibins of initializeObservation (initializeObservation.m:79-156): the
 * four edge vectors as observeTime bins them and the strides of the two
 * histogram variables; *n1, *n2 = bins of ibins(1), ibins(2). */
static void rtmc_mex_bins(rtmc_mex_keep* k, const mxArray* obs, const mxArray* ibins, int dirvar, rtmc_bins* out,
                          int32_t* nt, double* dt, mwSize* n1, mwSize* n2) {
    static const char* names[4] = {"binX", "binY", "binZ", "binPsi"};
    double ib[2];
    const mxArray* tv;
    size_t ntt;
    memset(out, 0, sizeof(*out));
    if (rtmc_mex_vector(k, ibins, ib, 2, "ibins") != 2) rtmc_mex_fail(k, "rtmc:arg", "ibins must hold two variables");
    for (int v = 0; v < 4; ++v) {
        const mxArray* e = rtmc_mex_get(obs, 0, names[v]);
        size_t ne;
        double* host;
        if (e == NULL) rtmc_mex_fail(k, "rtmc:arg", "obs.%s is missing", names[v]);
        ne = mxGetNumberOfElements(e);
        if (ne < 2) rtmc_mex_fail(k, "rtmc:arg", "obs.%s needs at least two edges", names[v]);
        host = (double*)mxMalloc(ne * sizeof(double));
        for (size_t i = 0; i < ne; ++i) host[i] = rtmc_mex_num(k, e, i, names[v]);
        out->n_edges[v] = (int32_t)ne;
        out->edges[v] = rtmc_mex_upload(k, host, ne);
        mxFree(host);
    }
    {
        const int a = (int)ib[0] - 1, b = (int)ib[1] - 1;
        if (a < 0 || a > 3 || b < 0 || b > 3 || a == b) rtmc_mex_fail(k, "rtmc:arg", "ibins must be two of 1..4");
        *n1 = (mwSize)(out->n_edges[a] - 1);
        *n2 = (mwSize)(out->n_edges[b] - 1);
        out->stride[a] = 1;
        out->stride[b] = (int64_t)*n1;
    }
    tv = rtmc_mex_get(obs, 0, "t");
    if (tv == NULL) rtmc_mex_fail(k, "rtmc:arg", "obs.t is missing");
    ntt = mxGetNumberOfElements(tv);
    *nt = (int32_t)ntt;
    *dt = 0.0;
    if (ntt > 1) {   /* mean(diff(t)) (radiativeTransfer.m:37) */
        double s = 0.0;
        for (size_t i = 1; i < ntt; ++i) s += rtmc_mex_num(k, tv, i, "obs.t") - rtmc_mex_num(k, tv, i - 1, "obs.t");
        *dt = s / (double)(ntt - 1);
    }
    out->time_stride = (int64_t)(*n1) * (int64_t)(*n2);
    out->mode_stride = out->time_stride * (int64_t)ntt;
    out->direction_variable = dirvar;
}

/* griddedInterpolant -> device table (MaterialClass.m:2459-2462) */
static void rtmc_mex_table(rtmc_mex_keep* k, const mxArray* gi, rtmc_table* out) {
    const mxArray* gv = rtmc_mex_get(gi, 0, "GridVectors");
    const mxArray* vals = rtmc_mex_get(gi, 0, "Values");
    const mxArray* x;
    size_t n;
    double* hx;
    double* hy;
    if (gv == NULL || !mxIsCell(gv) || vals == NULL) rtmc_mex_fail(k, "rtmc:arg", "invcdf must be a griddedInterpolant");
    x = mxGetCell(gv, 0);
    n = mxGetNumberOfElements(x);
    if (n < 2 || mxGetNumberOfElements(vals) != n) rtmc_mex_fail(k, "rtmc:arg", "malformed invcdf table");
    hx = (double*)mxMalloc(n * sizeof(double));
    hy = (double*)mxMalloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        hx[i] = rtmc_mex_num(k, x, i, "invcdf grid");
        hy[i] = rtmc_mex_num(k, vals, i, "invcdf values");
    }
    out->n = (int64_t)n;
    out->x = rtmc_mex_upload(k, hx, n);
    out->y = rtmc_mex_upload(k, hy, n);
    out->extrap = RTMC_EXTRAP_LINEAR;
    mxFree(hx);
    mxFree(hy);
}

#endif /* RTMC_MEX_UTIL_H */

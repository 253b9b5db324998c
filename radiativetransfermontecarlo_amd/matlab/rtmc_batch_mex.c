/*
 * rtmc_batch_mex.c -- MATLAB gateway of the batched forward solve and its
 * custom backward (rtmc_forward_batched, rtmc_adjoint_batched).  The
 * reference solves one material per radiativeTransfer call; a parameter study
 * or an inversion loop over materials that share the geometry and the
 * observation calls this once per sweep.  Materials are prepared natively
 * (rtmc_prepare_materials: MaterialClass.prepareSigma for isotropic media).
 *
 * Build in MATLAB (outside this repository's make):
 *   mex -R2018a rtmc_batch_mex.c -I<repo>/include -I/opt/rocm/include ...
 *       -L<repo>/radiativetransfermontecarlo_amd -lrtmc -L/opt/rocm/lib -lamdhip64
 *
 *   [E, counters, id] = rtmc_batch_mex('forward', geometry, sources, materials, obs, ibins, N, seeds[, options])
 *     sources    struct array with one source per material, or one struct for all
 *     materials  cell array of isotropic MaterialClass objects, prepared or not
 *     N, seeds   one value per material, or one value for all (seeds then seed + i - 1)
 *     options    struct: precision 'fp64' | 'fp32', hist_mode, autotune (fraction of the
 *                particles to tune the launch on; 0 = off), knot_threshold (default 1e-8,
 *                the reference's tables; 0 keeps every knot, which the adjoint assumes),
 *                lds_tol (radians; > 0: scattering angles from compressed tables held in
 *                LDS, within lds_tol of the full tables; the adjoint replays the same draws)
 *     E          cell array of uint32 [N1 N2 Nt 1+~acoustics], one per material
 *     counters   n x 4 double (scattering events, deposits, boundary interactions, removals)
 *     id         with three outputs, the batch is kept on the device for 'adjoint'
 *   grad = rtmc_batch_mex('adjoint', id, dLdE)
 *     dLdE       cell array of dL/dE (E's shape, single or double), one per material
 *     grad       n x 3: dL/dc, dL/dlc, dL/dQ (c = coefficients_of_variation(1)); acoustic media
 *   rtmc_batch_mex('free', id)
 *     releases a kept batch (all are released when the MEX file is cleared)
 */
#include <stdlib.h>

#include "rtmc_mex_util.h"

#define MAX_KEPT 64
#define N_THETA_DERIV 2049

typedef struct {
    rtmc_mex_keep dev;              /* device arrays of the batch (tables, edges, histograms) */
    void* mat_handle;               /* rtmc_prepare_materials */
    rtmc_batch batch;
    rtmc_batch_item* items;
    rtmc_material_params* params;
    mwSize dims[4];
} kept_batch;

static kept_batch* kept[MAX_KEPT];
/* the batch a 'forward' call is building: an error raised inside it returns to
   MATLAB without unwinding, so the next call releases it */
static kept_batch* pending;

static void free_batch(kept_batch* b) {
    if (!b) return;
    rtmc_mex_release(&b->dev);
    if (b->mat_handle) (void)rtmc_free_material(b->mat_handle);
    free(b->items);
    free(b->params);
    free(b);
}

static void at_exit(void) {
    for (int i = 0; i < MAX_KEPT; ++i) {
        free_batch(kept[i]);
        kept[i] = NULL;
    }
    free_batch(pending);
    pending = NULL;
}

/* Raises a MATLAB error after releasing a batch under construction. */
static void batch_fail(kept_batch* b, const char* what) {
    char msg[1024];
    snprintf(msg, sizeof(msg), "%s", what);
    free_batch(b);
    if (pending == b) pending = NULL;
    mexErrMsgIdAndTxt("rtmc:batch", "%s", msg);
}

/* MaterialClass -> rtmc_material_params (the properties prepareSigma reads, MaterialClass.m:2379-2463) */
static void params_from(rtmc_mex_keep* k, const mxArray* m, int d, double knot_threshold, rtmc_material_params* p) {
    char law[32], typ[32];
    const mxArray* cv = rtmc_mex_get(m, 0, "coefficients_of_variation");
    const mxArray* cc = rtmc_mex_get(m, 0, "correlation_coefficients");
    const mxArray* sp = rtmc_mex_get(m, 0, "SpectralParam");
    const mxArray* q = rtmc_mex_get(m, 0, "Q");
    double buf[3] = {0, 0, 0};
    size_t n;
    memset(p, 0, sizeof(*p));
    rtmc_mex_string(k, m, "type", typ, sizeof(typ));
    if (typ[0] != 0 && strcmp(typ, "isotropic") != 0)
        rtmc_mex_fail(k, "rtmc:arg", "native preparation covers isotropic media (calcSigmaIsotropic)");
    p->dimension = d;
    p->acoustics = rtmc_mex_scalar(k, m, "acoustics", 1.0) != 0.0;
    if (p->acoustics) {
        p->v[0] = rtmc_mex_scalar(k, m, "v", NAN);
    } else {
        p->v[0] = rtmc_mex_scalar(k, m, "vp", NAN);
        p->v[1] = rtmc_mex_scalar(k, m, "vs", NAN);
    }
    p->frequency = rtmc_mex_scalar(k, m, "Frequency", NAN);
    n = rtmc_mex_vector(k, cv, buf, 3, "coefficients_of_variation");
    for (size_t i = 0; i < n; ++i) p->cv[i] = buf[i];
    if (cc != NULL) {
        n = rtmc_mex_vector(k, cc, buf, 3, "correlation_coefficients");
        for (size_t i = 0; i < n; ++i) p->corr[i] = buf[i];
    }
    rtmc_mex_string(k, m, "SpectralLaw", law, sizeof(law));
    /* the law's name, or the display name a prepared material carries (MaterialClass.m:479-710) */
    if (strcmp(law, "exp") == 0) p->spectral_law = RTMC_PSDF_EXP;
    else if (strcmp(law, "power_law") == 0) p->spectral_law = RTMC_PSDF_POWER_LAW;
    else if (strcmp(law, "gaussian") == 0 || strcmp(law, "gauss") == 0) p->spectral_law = RTMC_PSDF_GAUSSIAN;
    else if (strcmp(law, "triangular") == 0) p->spectral_law = RTMC_PSDF_TRIANGULAR;
    else if (strcmp(law, "low_pass") == 0) p->spectral_law = RTMC_PSDF_LOW_PASS;
    else if (strcmp(law, "vonkarman") == 0) p->spectral_law = RTMC_PSDF_VON_KARMAN;
    else rtmc_mex_fail(k, "rtmc:arg", "SpectralLaw '%s' has no native preparation", law);
    p->lc = rtmc_mex_scalar(k, m, "CorrelationLength", NAN);
    p->nu = sp ? rtmc_mex_scalar(k, sp, "nu", 0.0) : 0.0;
    p->Q[0] = q ? rtmc_mex_num(k, q, 0, "Q") : INFINITY;
    p->Q[1] = (q && mxGetNumberOfElements(q) > 1) ? rtmc_mex_num(k, q, 1, "Q") : p->Q[0];
    p->knot_threshold = knot_threshold;
    p->n_theta = 0;
    p->rho = p->acoustics ? 0.0 : rtmc_mex_scalar(k, m, "rho", 0.0);
}

static void forward(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    const mxArray *geo = prhs[1], *srcs = prhs[2], *mats = prhs[3], *obs = prhs[4], *ibins = prhs[5];
    const mxArray *nv = prhs[6], *sv = prhs[7], *opt = nrhs > 8 ? prhs[8] : NULL;
    kept_batch* b = (kept_batch*)calloc(1, sizeof(kept_batch));
    rtmc_mex_keep* k;
    rtmc_material* dm;
    char prec[16];
    mwSize n1 = 0, n2 = 0;
    size_t n, total;
    double knot, tune, lds_tol;
    int slot = -1;
    if (!b) mexErrMsgIdAndTxt("rtmc:nomem", "out of host memory");
    pending = b;
    k = &b->dev;
    if (!mxIsCell(mats)) batch_fail(b, "materials must be a cell array of MaterialClass objects");
    n = mxGetNumberOfElements(mats);
    if (n == 0) batch_fail(b, "no materials");
    b->items = (rtmc_batch_item*)calloc(n, sizeof(rtmc_batch_item));
    b->params = (rtmc_material_params*)calloc(n, sizeof(rtmc_material_params));
    dm = (rtmc_material*)mxCalloc(n, sizeof(rtmc_material));
    if (!b->items || !b->params) batch_fail(b, "out of host memory");

    rtmc_mex_geometry(k, geo, &b->batch.geometry);
    rtmc_mex_bins(k, obs, ibins, RTMC_DIRVAR_DIRZ, &b->batch.bins, &b->batch.n_times, &b->batch.dt, &n1, &n2);
    knot = rtmc_mex_scalar(k, opt, "knot_threshold", 1e-8);
    lds_tol = rtmc_mex_scalar(k, opt, "lds_tol", 0.0);
    for (size_t i = 0; i < n; ++i) {
        params_from(k, mxGetCell(mats, (mwIndex)i), b->batch.geometry.dimension, knot, &b->params[i]);
        b->params[i].lds_tol = lds_tol;  /* > 0: compressed angle tables held in LDS */
    }
    rtmc_mex_check(k, rtmc_prepare_materials(b->params, (int32_t)n, dm, &b->mat_handle), "rtmc_prepare_materials");

    b->dims[0] = n1;
    b->dims[1] = n2;
    b->dims[2] = (mwSize)b->batch.n_times;
    total = (size_t)n1 * n2 * (size_t)b->batch.n_times;
    for (size_t i = 0; i < n; ++i) {
        rtmc_batch_item* it = &b->items[i];
        const size_t hist = total * (dm[i].acoustics ? 1 : 2);
        const mxArray* s = mxGetNumberOfElements(srcs) > 1 ? NULL : srcs;
        it->material = dm[i];
        if (s == NULL) {
            /* struct array: a one-element copy of source i */
            mxArray* one = mxCreateStructMatrix(1, 1, 0, NULL);
            for (int f = 0; f < mxGetNumberOfFields(srcs); ++f) {
                const char* name = mxGetFieldNameByNumber(srcs, f);
                const mxArray* v = mxGetFieldByNumber(srcs, (mwIndex)i, f);
                mxAddField(one, name);
                if (v) mxSetField(one, 0, name, mxDuplicateArray(v));
            }
            rtmc_mex_source(k, one, b->batch.geometry.dimension, &it->source);
            mxDestroyArray(one);
        } else {
            rtmc_mex_source(k, s, b->batch.geometry.dimension, &it->source);
        }
        it->n_particles = (int64_t)rtmc_mex_num(k, nv, mxGetNumberOfElements(nv) > 1 ? i : 0, "N");
        it->seed = (uint64_t)rtmc_mex_num(k, sv, mxGetNumberOfElements(sv) > 1 ? i : 0, "seeds") +
                   (mxGetNumberOfElements(sv) > 1 ? 0 : (uint64_t)i);
        it->particle_offset = 0;
        it->E = (uint32_t*)rtmc_mex_device(k, hist * sizeof(uint32_t));
        it->counters = (uint64_t*)rtmc_mex_device(k, RTMC_NCOUNTERS * sizeof(uint64_t));
        if (hipMemset(it->E, 0, hist * sizeof(uint32_t)) != hipSuccess ||
            hipMemset(it->counters, 0, RTMC_NCOUNTERS * sizeof(uint64_t)) != hipSuccess)
            batch_fail(b, "hipMemset failed");
    }
    b->batch.n_items = (int32_t)n;
    b->batch.items = b->items;
    b->batch.time_steps = (int32_t)rtmc_mex_scalar(k, mxGetCell(mats, 0), "timeSteps", 0.0);
    rtmc_mex_string(k, opt, "precision", prec, sizeof(prec));
    b->batch.precision = strcmp(prec, "fp32") == 0 ? RTMC_PREC_FP32_FLIGHT : RTMC_PREC_FP64;
    b->batch.hist_mode = (int32_t)rtmc_mex_scalar(k, opt, "hist_mode", RTMC_HIST_AUTO);
    tune = rtmc_mex_scalar(k, opt, "autotune", 0.0);
    if (tune > 0.0) {
        rtmc_tuning t;
        if (rtmc_autotune_batched(&b->batch, tune, &t, NULL) == RTMC_OK) {
            b->batch.block_size = t.block_size;
            b->batch.particles_per_thread = t.particles_per_thread;
            b->batch.hist_mode = t.hist_mode;
            b->batch.event_batch = t.event_batch;
        }
    }
    if (rtmc_forward_batched(&b->batch, NULL) != RTMC_OK) {
        char msg[1024];
        snprintf(msg, sizeof(msg), "rtmc_forward_batched failed: %s", rtmc_last_error());
        batch_fail(b, msg);
    }

    plhs[0] = mxCreateCellMatrix(1, n);
    if (nlhs > 1) plhs[1] = mxCreateDoubleMatrix(n, RTMC_NCOUNTERS, mxREAL);
    for (size_t i = 0; i < n; ++i) {
        mwSize dims[4] = {b->dims[0], b->dims[1], b->dims[2], b->items[i].material.acoustics ? 1 : 2};
        const size_t hist = total * dims[3];
        mxArray* e = mxCreateNumericArray(4, dims, mxUINT32_CLASS, mxREAL);
        uint64_t hc[RTMC_NCOUNTERS];
        if (hipMemcpy(mxGetUint32s(e), b->items[i].E, hist * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hc, b->items[i].counters, sizeof(hc), hipMemcpyDeviceToHost) != hipSuccess)
            batch_fail(b, "hipMemcpy to the host failed");
        mxSetCell(plhs[0], i, e);
        if (nlhs > 1)
            for (int c = 0; c < RTMC_NCOUNTERS; ++c) mxGetDoubles(plhs[1])[i + n * (size_t)c] = (double)hc[c];
    }
    if (nlhs > 2) {
        for (int i = 0; i < MAX_KEPT; ++i)
            if (!kept[i]) {
                slot = i;
                break;
            }
        if (slot < 0) batch_fail(b, "too many kept batches: free some with rtmc_batch_mex('free', id)");
        kept[slot] = b;
        plhs[2] = mxCreateDoubleScalar((double)(slot + 1));
    } else {
        free_batch(b);
    }
    pending = NULL;
}

static kept_batch* lookup(const mxArray* idv) {
    const double id = mxGetScalar(idv);
    const int s = (int)id - 1;
    if (s < 0 || s >= MAX_KEPT || (double)(s + 1) != id || !kept[s])
        mexErrMsgIdAndTxt("rtmc:batch", "no kept batch with this id");
    return kept[s];
}

static void adjoint(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    kept_batch* b;
    const mxArray* dl;
    rtmc_mex_keep tmp;
    rtmc_adjoint_item* a;
    double* g;
    double* dS;
    size_t n, total;
    (void)nlhs;
    if (nrhs != 3) mexErrMsgIdAndTxt("rtmc:usage", "grad = rtmc_batch_mex('adjoint', id, dLdE)");
    b = lookup(prhs[1]);
    dl = prhs[2];
    n = (size_t)b->batch.n_items;
    if (!mxIsCell(dl) || mxGetNumberOfElements(dl) != n)
        mexErrMsgIdAndTxt("rtmc:arg", "dLdE must be a cell array with one array per material");
    tmp.n = 0;
    tmp.nh = 0;
    total = (size_t)b->dims[0] * b->dims[1] * b->dims[2];
    a = (rtmc_adjoint_item*)mxCalloc(n, sizeof(rtmc_adjoint_item));
    g = (double*)mxCalloc(n * N_THETA_DERIV, sizeof(double));
    dS = (double*)mxCalloc(n, sizeof(double));
    /* d log sigma / d lc on linspace(0, pi, N_THETA_DERIV) and d log Sigma / d lc */
    rtmc_mex_check(&tmp, rtmc_lc_derivatives(b->params, (int32_t)n, N_THETA_DERIV, 1e-4, g, dS), "rtmc_lc_derivatives");
    for (size_t i = 0; i < n; ++i) {
        const mxArray* w = mxGetCell(dl, i);
        float* hw;
        if (!b->params[i].acoustics) rtmc_mex_fail(&tmp, "rtmc:arg", "the adjoint covers acoustic media");
        if (w == NULL || mxGetNumberOfElements(w) != total)
            rtmc_mex_fail(&tmp, "rtmc:arg", "dLdE{%zu} must have the %zu elements of E{%zu}", i + 1, total, i + 1);
        hw = (float*)mxMalloc(total * sizeof(float));
        if (mxIsSingle(w)) memcpy(hw, mxGetSingles(w), total * sizeof(float));
        else if (mxIsDouble(w))
            for (size_t j = 0; j < total; ++j) hw[j] = (float)mxGetDoubles(w)[j];
        else rtmc_mex_fail(&tmp, "rtmc:arg", "dLdE{%zu} must be single or double", i + 1);
        a[i].weight = (const float*)rtmc_mex_device(&tmp, total * sizeof(float));
        if (hipMemcpy((void*)a[i].weight, hw, total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            rtmc_mex_fail(&tmp, "rtmc:hip", "hipMemcpy to the device failed");
        mxFree(hw);
        a[i].grad = (double*)rtmc_mex_device(&tmp, 3 * sizeof(double));
        if (hipMemset(a[i].grad, 0, 3 * sizeof(double)) != hipSuccess) rtmc_mex_fail(&tmp, "rtmc:hip", "hipMemset failed");
        a[i].dhist = NULL;
        a[i].cv = b->params[i].cv[0];
        a[i].dlogSigma_dlc = dS[i];
        a[i].dlogsigma_dlc = rtmc_mex_upload(&tmp, g + i * N_THETA_DERIV, N_THETA_DERIV);
        a[i].n_dlogsigma = N_THETA_DERIV;
    }
    rtmc_mex_check(&tmp, rtmc_adjoint_batched(&b->batch, a, NULL), "rtmc_adjoint_batched");
    plhs[0] = mxCreateDoubleMatrix(n, 3, mxREAL);
    for (size_t i = 0; i < n; ++i) {
        double h[3];
        if (hipMemcpy(h, a[i].grad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            rtmc_mex_fail(&tmp, "rtmc:hip", "hipMemcpy to the host failed");
        for (int c = 0; c < 3; ++c) mxGetDoubles(plhs[0])[i + n * (size_t)c] = h[c];
    }
    rtmc_mex_release(&tmp);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    char cmd[16];
    mexAtExit(at_exit);
    free_batch(pending);   /* left by a 'forward' call that raised an error */
    pending = NULL;
    if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], cmd, sizeof(cmd)) != 0)
        mexErrMsgIdAndTxt("rtmc:usage", "rtmc_batch_mex('forward' | 'adjoint' | 'free', ...)");
    if (rtmc_abi_version() != RTMC_ABI_VERSION)
        mexErrMsgIdAndTxt("rtmc:abi", "librtmc ABI %d, this gateway was built for %d", rtmc_abi_version(), RTMC_ABI_VERSION);
    if (strcmp(cmd, "forward") == 0) {
        if (nrhs < 8 || nrhs > 9 || nlhs > 3)
            mexErrMsgIdAndTxt("rtmc:usage", "[E, counters, id] = rtmc_batch_mex('forward', geometry, sources, "
                                            "materials, obs, ibins, N, seeds[, options])");
        forward(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "adjoint") == 0) {
        adjoint(nlhs, plhs, nrhs, prhs);
    } else if (strcmp(cmd, "free") == 0) {
        kept_batch* b;
        if (nrhs != 2) mexErrMsgIdAndTxt("rtmc:usage", "rtmc_batch_mex('free', id)");
        b = lookup(prhs[1]);
        for (int i = 0; i < MAX_KEPT; ++i)
            if (kept[i] == b) kept[i] = NULL;
        free_batch(b);
    } else {
        mexErrMsgIdAndTxt("rtmc:usage", "unknown command '%s'", cmd);
    }
}

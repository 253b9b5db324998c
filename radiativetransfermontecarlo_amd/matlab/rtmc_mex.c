/*
 * rtmc_mex.c -- MATLAB gateway of rtmc_forward: the packet x time loop of
 * radiativeTransfer.m:40-158 in one call on the GPU.
 *
 * Build in MATLAB (outside this repository's make):
 *   mex -R2018a rtmc_mex.c -I<repo>/include -I/opt/rocm/include ...
 *       -L<repo>/radiativetransfermontecarlo_amd -lrtmc -L/opt/rocm/lib -lamdhip64
 *
 * In radiativeTransfer.m, after initializeObservation (:27-28) and
 * prepareSigma (:31), lines 40-158 become
 *   [energy, counters] = rtmc_mex(geometry, source, material, obs, ibins, Np*Npk, seed);
 * and the normalisation (:160-191) is unchanged.
 *
 * Inputs
 *   geometry  struct: dimension, frame, bnd(i).dir/.val/.type
 *   source    struct of initializeParticle.m (point or plane; radial law as a griddedInterpolant)
 *   material  MaterialClass after prepareSigma: acoustics, v|vp,vs, meanFreeTime, P2P, S2S,
 *             Q, Frequency, invcdf, timeSteps, rho (elastic reflective boundaries)
 *   obs, ibins  outputs 1 and 4 of initializeObservation
 *   N         particles;  seed  Philox key (any integer below 2^53)
 *   options   (optional) struct: precision 'fp64' (default) | 'fp32' (FP32 free flight,
 *             FP64 time), hist_mode 0..3 (RTMC_HIST_*), direction_variable 'dirz' (default,
 *             observeTime.m) | 'position' (README.md:209), lds_tol (radians; > 0: draw the
 *             scattering angles from compressed copies of invcdf held in LDS, within lds_tol
 *             of the full tables, rtmc_attach_lds_tables; dirz only)
 * Outputs
 *   energy    uint32 [N1 N2 Nt 1+~acoustics], the energy array initializeObservation
 *             allocates, filled
 *   counters  1 x 4 double: scattering events, deposits, boundary interactions, removals
 */
#include "rtmc_mex_util.h"

/* MaterialClass after prepareSigma -> rtmc_material (MaterialClass.m:2379-2442) */
static void material_from(rtmc_mex_keep* k, const mxArray* mat, int with_zoeppritz, rtmc_material* m) {
    const mxArray* inv = rtmc_mex_get(mat, 0, "invcdf");
    const mxArray* mft = rtmc_mex_get(mat, 0, "meanFreeTime");
    const mxArray* q = rtmc_mex_get(mat, 0, "Q");
    memset(m, 0, sizeof(*m));
    m->acoustics = rtmc_mex_scalar(k, mat, "acoustics", 1.0) != 0.0;
    if (inv == NULL || mft == NULL) rtmc_mex_fail(k, "rtmc:arg", "material must be prepared (MaterialClass.prepareSigma)");
    if (m->acoustics) {
        const double v = rtmc_mex_scalar(k, mat, "vp", rtmc_mex_scalar(k, mat, "v", NAN));
        m->vp = v;
        m->vs = 0.0;
        m->mft[0] = m->mft[1] = rtmc_mex_num(k, mft, 0, "meanFreeTime");
        m->P2P = m->S2S = 1.0;
        rtmc_mex_table(k, inv, &m->invcdf[0]);
    } else {
        /* invcdf{i,j}: from mode i to mode j, stored column-major in the 2 x 2 cell */
        static const int cell_of[4] = {0, 2, 1, 3};   /* P->P, P->S, S->P, S->S */
        m->vp = rtmc_mex_scalar(k, mat, "vp", NAN);
        m->vs = rtmc_mex_scalar(k, mat, "vs", NAN);
        m->mft[0] = rtmc_mex_num(k, mft, 0, "meanFreeTime");
        m->mft[1] = mxGetNumberOfElements(mft) > 1 ? rtmc_mex_num(k, mft, 1, "meanFreeTime") : m->mft[0];
        m->P2P = rtmc_mex_scalar(k, mat, "P2P", NAN);
        m->S2S = rtmc_mex_scalar(k, mat, "S2S", NAN);
        if (!mxIsCell(inv) || mxGetNumberOfElements(inv) != 4)
            rtmc_mex_fail(k, "rtmc:arg", "an elastic material's invcdf must be a 2 x 2 cell");
        for (int c = 0; c < 4; ++c) rtmc_mex_table(k, mxGetCell(inv, cell_of[c]), &m->invcdf[c]);
    }
    m->QP = q ? rtmc_mex_num(k, q, 0, "Q") : INFINITY;
    m->QS = (q && !m->acoustics && mxGetNumberOfElements(q) > 1) ? rtmc_mex_num(k, q, 1, "Q") : m->QP;
    m->frequency = rtmc_mex_scalar(k, mat, "Frequency", NAN);
    if (with_zoeppritz && !m->acoustics) {
        /* [out, ~] = MaterialClass.Zoeppritz(mat): solid under air, 181 angles */
        enum { NZ = 181 };
        double deg[NZ], rpp[NZ], rss[NZ];
        const double rho = rtmc_mex_scalar(k, mat, "rho", NAN);
        if (!(rho > 0.0))
            rtmc_mex_fail(k, "rtmc:arg", "Please define the density in MaterialClass before calling MaterialInterface.");
        rtmc_mex_check(k, rtmc_zoeppritz_air(m->vp, m->vs, rho, NZ, deg, rpp, rss), "rtmc_zoeppritz_air");
        m->n_zoe = NZ;
        m->zoe_deg = rtmc_mex_upload(k, deg, NZ);
        m->zoe_Rpp = rtmc_mex_upload(k, rpp, NZ);
        m->zoe_Rsvsv = rtmc_mex_upload(k, rss, NZ);
    }
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    rtmc_mex_keep keep;
    rtmc_problem p;
    const mxArray* opt = nrhs > 7 ? prhs[7] : NULL;
    char prec[16], dv[16];
    int reflective = 0;
    mwSize n1 = 0, n2 = 0, dims[4];
    size_t total;
    uint32_t* dE;
    uint64_t* dC;
    uint64_t hc[RTMC_NCOUNTERS];
    keep.n = 0;
    keep.nh = 0;
    if (nrhs < 7 || nrhs > 8 || nlhs > 2)
        mexErrMsgIdAndTxt("rtmc:usage", "[energy, counters] = rtmc_mex(geometry, source, material, obs, ibins, N, seed[, options])");
    if (rtmc_abi_version() != RTMC_ABI_VERSION)
        mexErrMsgIdAndTxt("rtmc:abi", "librtmc ABI %d, this gateway was built for %d", rtmc_abi_version(), RTMC_ABI_VERSION);
    memset(&p, 0, sizeof(p));

    rtmc_mex_geometry(&keep, prhs[0], &p.geometry);
    for (int i = 0; i < p.geometry.n_bnd; ++i) reflective |= p.geometry.bnd[i].type == RTMC_BND_REFLECTIVE;
    rtmc_mex_source(&keep, prhs[1], p.geometry.dimension, &p.source);
    material_from(&keep, prhs[2], reflective, &p.material);
    rtmc_mex_string(&keep, opt, "direction_variable", dv, sizeof(dv));
    rtmc_mex_bins(&keep, prhs[3], prhs[4], strcmp(dv, "position") == 0 ? RTMC_DIRVAR_POSITION : RTMC_DIRVAR_DIRZ,
                  &p.bins, &p.n_times, &p.dt, &n1, &n2);
    {
        const double lds_tol = rtmc_mex_scalar(&keep, opt, "lds_tol", 0.0);
        if (lds_tol > 0.0) {
            void* h = NULL;
            rtmc_mex_check(&keep, rtmc_attach_lds_tables(&p.material, 1, lds_tol, 0, &h, NULL), "rtmc_attach_lds_tables");
            keep.h[keep.nh++] = h;
        }
    }

    p.n_particles = (int64_t)rtmc_mex_num(&keep, prhs[5], 0, "N");
    p.seed = (uint64_t)rtmc_mex_num(&keep, prhs[6], 0, "seed");
    p.particle_offset = 0;
    p.time_steps = (int32_t)rtmc_mex_scalar(&keep, prhs[2], "timeSteps", 0.0);
    rtmc_mex_string(&keep, opt, "precision", prec, sizeof(prec));
    p.precision = strcmp(prec, "fp32") == 0 ? RTMC_PREC_FP32_FLIGHT : RTMC_PREC_FP64;
    p.hist_mode = (int32_t)rtmc_mex_scalar(&keep, opt, "hist_mode", RTMC_HIST_AUTO);

    /* energy = zeros(N1, N2, Nt, 1+~acoustics, 'uint32') (initializeObservation.m:133) */
    dims[0] = n1;
    dims[1] = n2;
    dims[2] = (mwSize)p.n_times;
    dims[3] = p.material.acoustics ? 1 : 2;
    total = (size_t)n1 * n2 * (size_t)p.n_times * dims[3];
    dE = (uint32_t*)rtmc_mex_device(&keep, total * sizeof(uint32_t));
    dC = (uint64_t*)rtmc_mex_device(&keep, sizeof(hc));
    if (hipMemset(dE, 0, total * sizeof(uint32_t)) != hipSuccess || hipMemset(dC, 0, sizeof(hc)) != hipSuccess)
        rtmc_mex_fail(&keep, "rtmc:hip", "hipMemset failed");

    rtmc_mex_check(&keep, rtmc_forward(&p, dE, dC, NULL), "rtmc_forward");

    plhs[0] = mxCreateNumericArray(4, dims, mxUINT32_CLASS, mxREAL);
    if (hipMemcpy(mxGetUint32s(plhs[0]), dE, total * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hc, dC, sizeof(hc), hipMemcpyDeviceToHost) != hipSuccess)
        rtmc_mex_fail(&keep, "rtmc:hip", "hipMemcpy to the host failed");
    if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(1, RTMC_NCOUNTERS, mxREAL);
        for (int i = 0; i < RTMC_NCOUNTERS; ++i) mxGetDoubles(plhs[1])[i] = (double)hc[i];
    }
    rtmc_mex_release(&keep);
}

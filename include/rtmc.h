/*
 * rtmc.h -- C ABI of librtmc, the MI355X (gfx950) particle-tracing core of
 * radiativeTransfer(geometry, source, material, observation).
 *
 * Every entry point replaces one function (or loop) of the reference MATLAB
 * code, cited below as file:line of cottereau/RadiativeTransferMonteCarlo.
 * The reference has no FFI of its own: these are the symbols a MEX gateway
 * (mexFunction) or a Python ctypes binding calls.
 *
 * Conventions
 *  - Plain C types only. Bulk arrays are DEVICE pointers (hipMalloc'ed or
 *    taken from a device tensor); the small descriptor structs live in host
 *    memory and are passed by pointer.
 *  - Arrays of N x 3 values are stored column-major exactly as MATLAB stores
 *    an N x 3 double matrix: x[0..N-1] = x(:,1), x[N..2N-1] = x(:,2), ...
 *  - Histograms are uint32 counts, column-major over their bin axes, then
 *    time, then mode (mode 1 = S waves), as E in initializeObservation.m:133.
 *  - Every function returns RTMC_OK (0) or a negative status; the message of
 *    the last failure on the calling thread is returned by rtmc_last_error().
 *    No C++ exception crosses the ABI.
 *  - `stream` is a hipStream_t (NULL = the default stream). Calls are
 *    stream-ordered; functions that must read a device value back (the
 *    mean particle time, an overflow flag) synchronise that stream.
 */
#ifndef RTMC_H
#define RTMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTMC_ABI_VERSION 6

/* status codes */
#define RTMC_OK 0
#define RTMC_ERR_ARG (-1)         /* invalid argument / unsupported combination */
#define RTMC_ERR_HIP (-2)         /* a HIP runtime call failed */
#define RTMC_ERR_DRAWS (-3)       /* replay stream shorter than the work needs */
#define RTMC_ERR_NOMEM (-4)       /* device allocation failed */
#define RTMC_ERR_TABLE (-5)       /* malformed interpolation table */

#define RTMC_MAX_BND 8

/* geometry.frame (initializeObservation.m:30-41) */
#define RTMC_FRAME_CARTESIAN 0
#define RTMC_FRAME_CYLINDRICAL 1
#define RTMC_FRAME_SPHERICAL 2

/* geometry.bnd(i).type (propagateParticleSmallDt.m:152-165) */
#define RTMC_BND_REFLECTIVE 0
#define RTMC_BND_ABSORBING 1

/* source.type / source.direction / source.aperture (initializeParticle.m) */
#define RTMC_SRC_POINT 0
#define RTMC_SRC_PLANE 1
#define RTMC_DIR_UNIFORM 0
#define RTMC_DIR_OUTGOING 1
#define RTMC_DIR_UPPER 2
#define RTMC_APERTURE_RECTANGLE 0
#define RTMC_APERTURE_CIRCLE 1

/* interpolation-table extrapolation rule */
#define RTMC_EXTRAP_LINEAR 0      /* griddedInterpolant default (MaterialClass.m:2462) */
#define RTMC_EXTRAP_NEAREST 1     /* griddedInterpolant(...,'linear','nearest') (initializeParticle.m:201) */

/* random-number source of a call */
#define RTMC_DRAWS_REPLAY 0       /* recorded variates, one slot per particle per step */
#define RTMC_DRAWS_PHILOX 1       /* Philox4x32-10 counter-based generator */

/* arithmetic of the fused propagate kernel */
#define RTMC_PREC_FP64 0
#define RTMC_PREC_FP32_FLIGHT 1   /* FP32 direction/flight arithmetic, FP64 time */

/* One boundary: geometry.bnd(i) = struct('dir',dir,'val',val,'type',type).
 * dir 1..3 = plane normal to that axis at coordinate val; dir 4 = cylinder
 * of radius val around the z axis (propagateParticleSmallDt.m:167-581). */
typedef struct {
    int32_t dir;
    int32_t type;               /* RTMC_BND_* */
    double val;
} rtmc_boundary;

typedef struct {
    int32_t dimension;          /* 2 or 3 */
    int32_t frame;              /* RTMC_FRAME_* */
    int32_t n_bnd;              /* 0 .. RTMC_MAX_BND */
    int32_t pad_;
    rtmc_boundary bnd[RTMC_MAX_BND];
} rtmc_geometry;

/* A piecewise-linear table y(x) on strictly increasing knots x: the
 * griddedInterpolant(x, y) of MaterialClass.prepareSigmaOne
 * (MaterialClass.m:2459-2462) or of initializeParticle.m:201. */
typedef struct {
    int64_t n;                  /* number of knots (>= 2), 0 = table absent */
    const double* x;            /* device, length n */
    const double* y;            /* device, length n */
    int32_t extrap;             /* RTMC_EXTRAP_* */
    int32_t pad_;
} rtmc_table;

/* One knot of a compressed inverse-CDF table (rtmc_compress_table):
 * theta(q) = theta + (q - u) * slope for u <= q < the next knot's u. */
typedef struct {
    double u;                   /* cdf value of the knot (one of the table's own x) */
    float theta;                /* the table's angle there */
    float slope;                /* d theta / d u up to the next knot */
} rtmc_lds_knot;

/* No counterpart in the reference: an inverse-CDF table in the form the
 * tracing kernel holds in LDS.  m of the table's own knots, chosen so that
 * |theta(q) - table(q)| <= tol for every q (plus float rounding of theta and
 * slope, <= 4e-7); the first and the last two knots of the table are kept,
 * so the linear extrapolation beyond them is the table's own. */
typedef struct {
    int32_t m;                  /* knots (0 = none: the kernel reads the full table from global memory) */
    int32_t extrap;             /* the table's RTMC_EXTRAP_* (NEAREST: q is clamped to the end knots) */
    double tol;                 /* the bound above, in radians */
    const rtmc_lds_knot* knots; /* device, m */
} rtmc_lds_table;

/* Prepared material: the fields of MaterialClass that the particle kernels
 * read after MaterialClass.prepareSigma (MaterialClass.m:2379-2442). */
typedef struct {
    int32_t acoustics;          /* 1 acoustic, 0 elastic */
    int32_t pad_;
    double vp;                  /* acoustic: mat.v (prepareSigma sets vp = v) */
    double vs;
    double mft[2];              /* meanFreeTime of P and S (acoustic: both 1/Sigma); +Inf = no scattering */
    double P2P;                 /* Sigma(1,1)/sum(Sigma(1,:)) */
    double S2S;                 /* Sigma(2,2)/sum(Sigma(2,:)) */
    double QP;                  /* +Inf = no intrinsic attenuation */
    double QS;
    double frequency;           /* mat.Frequency(1) */
    rtmc_table invcdf[4];       /* [0] P->P (acoustic: the only table), [1] P->S, [2] S->P, [3] S->S */
    /* Solid/air energy coefficients on the Zoeppritz angle grid
     * (MaterialClass.m:1786-1830, ZoeppritzFluid.m), used by elastic
     * reflective boundaries only. */
    int64_t n_zoe;              /* 0 or number of angles (181) */
    const double* zoe_deg;      /* device: incidence angles in degrees */
    const double* zoe_Rpp;      /* device: E_Rpp */
    const double* zoe_Rsvsv;    /* device: E_Rsvsv */
    /* Optional compressed copies of invcdf for the tracing kernel
     * (rtmc_attach_lds_tables); m = 0 leaves every draw on invcdf itself. */
    rtmc_lds_table lds[4];
} rtmc_material;

/* Source description (initializeParticle.m:3-160, README.md:180-198). */
typedef struct {
    int32_t type;               /* RTMC_SRC_* */
    int32_t direction;          /* point: RTMC_DIR_* */
    int32_t aperture;           /* plane: RTMC_APERTURE_* */
    int32_t polarization;       /* 0 = P (default), 1 = S */
    double position[3];
    double lambda;
    double normal[3];           /* plane: unit propagation direction n */
    double basis_u[3];          /* plane: first column of null(n) */
    double basis_v[3];          /* plane: second column of null(n) */
    double extent[2];           /* plane: rectangle sides [L1 L2] or circle radius in [0] */
    rtmc_table radial;          /* point: user radial inverse CDF (n = 0: Gaussian radius) */
} rtmc_source;

/* Particle packet, structure of arrays (initializeParticle.m:163-170). */
typedef struct {
    int64_t n;
    int32_t dimension;
    int32_t pad_;
    double* x;                  /* device, 3n, column-major */
    double* dir;                /* device, 3n */
    double* perp;               /* device, 3n */
    uint8_t* p;                 /* device, n: 1 = P wave */
    uint8_t* alive;             /* device, n */
    double* t;                  /* device, n */
} rtmc_particles;

/* Random variates of one call.
 * REPLAY: replay[(step * n_slots + slot) * n + i] is the variate of particle
 * i at step `step` for draw site `slot` (the slot tables are listed with each
 * entry point). n_steps is the number of steps the buffer holds.
 * PHILOX: variate = uniform from philox4x32_10(counter, key) with
 * counter = {i & 0xffffffff, (i >> 32) | slot << 16, step, stream},
 * key = {seed & 0xffffffff, seed >> 32}, uniform = (bits53 + 1) * 2^-53. */
typedef struct {
    int32_t kind;               /* RTMC_DRAWS_* */
    int32_t n_slots;            /* REPLAY: slots per step (must equal the entry point's count) */
    const double* replay;       /* REPLAY: device buffer */
    int64_t n_steps;            /* REPLAY: steps held */
    uint64_t seed;              /* PHILOX */
    uint64_t stream;            /* PHILOX: call identifier (low 32 bits used) */
} rtmc_draws;

/* rtmc_bins.direction_variable: what variable 3 measures.
 * DIRZ: the direction's z component (observeTime.m:24-35), the reference's code;
 * POSITION: the angle between the direction and the position vector, as
 * README.md:209 documents it (cos of it in 3-D, the angle in 2-D; a particle
 * at the origin has none and is not counted). */
#define RTMC_DIRVAR_DIRZ 0
#define RTMC_DIRVAR_POSITION 1

/* Observation binning.  Variable v in {0: x / radius, 1: y / azimuth,
 * 2: z / sin(elevation), 3: direction (dir_z in 3-D, acos(dir_z) in 2-D, or
 * see RTMC_DIRVAR_POSITION)} as computed by observeTime.m:15-45.  A particle is counted when every
 * variable falls inside its edges (histcounts rule: e_k <= v < e_k+1, the
 * last bin closed); bin k of variable v adds k * stride[v] to the index.
 * For the reference's two-variable histograms (initializeObservation.m:79-130)
 * the two binned variables have strides 1 and N1, the others one bin each.
 * In 2-D variable 2 does not exist and is ignored. */
typedef struct {
    int32_t n_edges[4];         /* >= 2 each */
    const double* edges[4];     /* device, non-decreasing */
    int64_t stride[4];
    int64_t time_stride;        /* index step of one observation time */
    int64_t mode_stride;        /* index step of mode 1 (S); unused for acoustics */
    int32_t direction_variable; /* RTMC_DIRVAR_* */
    int32_t pad_;
} rtmc_bins;

/* Histogram placement of the fused kernel (rtmc_problem.hist_mode).
 * AUTO: LDS when the whole histogram (+ edges) fits in a work-group's LDS,
 * otherwise TMAJOR (GLOBAL if its scratch cannot be allocated).
 * TMAJOR deposits into a time-major scratch copy (time fastest), where a
 * particle that stays in one bin over consecutive observation times costs
 * one atomic for up to three of them (21-bit counters, one scratch copy per
 * 2^21 - 1 particles, while all copies fit in 4 GiB) or for two (32-bit
 * counters, otherwise), and adds it into E with a tiled transpose; the
 * scratch comes from the device's default memory pool (kept by the pool for
 * the next call). */
#define RTMC_HIST_AUTO 0
#define RTMC_HIST_GLOBAL 1
#define RTMC_HIST_LDS 2
#define RTMC_HIST_TMAJOR 3

/* Whole forward problem for the fused kernel (radiativeTransfer.m:20-158). */
typedef struct {
    rtmc_geometry geometry;
    rtmc_source source;
    rtmc_material material;
    rtmc_bins bins;
    int64_t n_particles;        /* particles of this call */
    uint64_t particle_offset;   /* global id of the first particle (chunking / sharding) */
    uint64_t seed;
    int32_t n_times;            /* Nt, including t = 0 which receives no deposit */
    int32_t time_steps;         /* material.timeSteps: 0 sub-stepped (default), 1 event-driven */
    double dt;                  /* mean(diff(t)) (radiativeTransfer.m:37) */
    int32_t precision;          /* RTMC_PREC_* */
    int32_t block_size;         /* 0 = library default */
    int32_t particles_per_thread; /* 0 = library default */
    int32_t hist_mode;          /* RTMC_HIST_*: 0 = auto, 1 = global atomics, 2 = LDS-privatised,
                                   3 = time-major scratch (runs of up to three or two times per atomic) */
} rtmc_problem;

/* Information returned by rtmc_propagate_smalldt about its sub-steps. */
typedef struct {
    double T_eff;               /* T - mean(P.t) (propagateParticleSmallDt.m:4) */
    double dt;                  /* sub-step (propagateParticleSmallDt.m:5-7) */
    int64_t n_sub;              /* number of sub-steps */
    double pkill_p;             /* 1 - exp(-omega dt / QP) */
    double pkill_s;
} rtmc_substep_info;

/* Counters accumulated by rtmc_forward (device uint64[RTMC_NCOUNTERS]). */
#define RTMC_CNT_SCATTER 0      /* scattering events */
#define RTMC_CNT_DEPOSIT 1      /* histogram increments */
#define RTMC_CNT_BOUNDARY 2     /* boundary interactions */
#define RTMC_CNT_KILLED 3       /* particles removed by Q or absorbing boundaries */
#define RTMC_NCOUNTERS 4

/* ------------------------------------------------------------------ */

/* Message of the last failure on this thread ("" if none). */
const char* rtmc_last_error(void);

/* RTMC_ABI_VERSION of the loaded library. */
int rtmc_abi_version(void);

/* Replaces initializeParticle(N, d, acoustics, source)
 * (initializeParticle.m:1-171, local inverseCDF :175-203).
 * Fills every array of P (P->n particles, P->dimension = d).
 * Draw slots (one step): 0 randn radius (plane: randn depth), 1 u phi0,
 * 2 u phi, 3 u theta0, 4 u theta, 5 u theta ('upper'), 6 u phi ('upper'),
 * 7 u alpha; plane source: 1 u transverse 1 (circle: radius), 2 u transverse 2
 * (circle: angle), 4 u polarisation angle.  n_slots = 8 (REPLAY).  Slot 0
 * under PHILOX is a standard normal made by Box-Muller from the slot-0 and
 * slot-8 uniforms. */
int rtmc_init_particles(const rtmc_source* source, int32_t acoustics,
                        const rtmc_draws* draws, rtmc_particles* P, void* stream);

/* Replaces propagateParticleSmallDt(mat, geometry, P, T)
 * (propagateParticleSmallDt.m:1-647, with MaterialClass.MaterialInterface
 * MaterialClass.m:1572-1756 for elastic plane boundaries).
 * T_eff: NaN = compute T - mean(P.t) on the device (as :4), otherwise used
 * as given (replay of a recorded MATLAB value).  `info` may be NULL.
 * Draw slots per sub-step: 0 u kill, 1 scatter (REPLAY: the poissrnd count;
 * PHILOX: uniform compared with 1-exp(-dt/mft)), 2 azimuth (3-D uniform;
 * 2-D: the randi([0 1]) bit), 3 u theta, 4 u P2P, 5 u S2S, then for
 * boundary b: 6+2b first boundary draw (P: reflect-or-convert; S: SV/SH
 * split), 7+2b second (SV: reflect-or-convert).  n_slots = 6 + 2 n_bnd. */
int rtmc_propagate_smalldt(const rtmc_material* mat, const rtmc_geometry* geometry,
                           rtmc_particles* P, double T, double T_eff,
                           const rtmc_draws* draws, rtmc_substep_info* info,
                           void* stream);

/* Replaces propagateParticle(mat, P, T) (propagateParticle.m:1-96).
 * Each particle loops until it reaches T; draw step = its loop iteration.
 * Draw slots: 0 event count (REPLAY: the poissrnd count; PHILOX: Poisson
 * sample by inversion of the slot-0 uniform), 1 minimum of the count's
 * uniforms (timeNextJump :86-96; PHILOX draws it as 1-u^(1/count)),
 * 2 azimuth, 3 u theta, 4 u P2P, 5 u S2S.  n_slots = 6 (REPLAY). */
int rtmc_propagate_events(const rtmc_material* mat, rtmc_particles* P, double T,
                          const rtmc_draws* draws, void* stream);

/* Replaces E(:,:,it,:) = E(:,:,it,:) + observeTime(...)
 * (observeTime.m:1-146 and radiativeTransfer.m:75-76): adds the histogram of
 * the alive particles of P into E at time index `time_index`. */
int rtmc_observe(const rtmc_geometry* geometry, const rtmc_bins* bins,
                 int32_t acoustics, const rtmc_particles* P,
                 uint32_t* E, int64_t time_index, void* stream);

/* Replaces the packet x time loop of radiativeTransfer
 * (radiativeTransfer.m:40-158): births, propagates and observes
 * prob->n_particles particles in one persistent kernel with Philox draws,
 * adding their counts into E (device, caller-zeroed, see rtmc_bins).
 * time_steps 0 samples exactly the sub-step process of
 * propagateParticleSmallDt (scattering and Q removal tested once per
 * sub-step), event to event; time_steps 1 is the event-driven process of
 * propagateParticle.  counters: NULL or device uint64[RTMC_NCOUNTERS]. */
int rtmc_forward(const rtmc_problem* prob, uint32_t* E, uint64_t* counters,
                 void* stream);

/* Bytes of device scratch rtmc_forward_noalloc may use per GPU. */
#define RTMC_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference.  rtmc_forward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, same histogram and
 * counters as rtmc_forward, and no hipMalloc / hipFree / pool allocation.
 * The call's device scratch (problem record, particle counters,
 * inverse-CDF guide, time-major histogram scratch) is carved from a
 * library-held array of RTMC_NOALLOC_SCRATCH_BYTES on each GPU, which the
 * GPU keeps until the process ends; its contents are cleared by kernels at
 * every call.  What does not fit is done without it, as rtmc_forward does
 * when memory is short (guide: search of the whole table; time-major
 * scratch: pairs from even times only, then direct atomics into E), except
 * an explicit hist_mode RTMC_HIST_TMAJOR whose scratch does not fit, which
 * returns RTMC_ERR_ARG.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams or threads) race and must be
 * serialised by the caller. */
int rtmc_forward_noalloc(const rtmc_problem* prob, uint32_t* E, uint64_t* counters,
                         void* stream);

/* ------------------------------------------------------------------ */
/* Batched launches: many materials and sources with their own particle
 * counts in one persistent kernel.  The reference solves one material per
 * radiativeTransfer call (radiativeTransfer.m:1-207); a batch replaces a loop
 * of such calls over materials that share the geometry, the observation bins
 * and the time axis. */

/* One material + source of a batch. */
typedef struct {
    rtmc_material material;
    rtmc_source source;
    int64_t n_particles;
    uint64_t particle_offset;   /* global id of the first particle */
    uint64_t seed;              /* Philox key; same seed + ids = same trajectories */
    uint32_t* E;                /* device histogram (rtmc_bins layout); counts are added */
    uint64_t* counters;         /* device uint64[RTMC_NCOUNTERS] (added to) or NULL */
} rtmc_batch_item;

/* What the items of a batch share, and the launch configuration. */
typedef struct {
    rtmc_geometry geometry;
    rtmc_bins bins;
    int32_t n_times;            /* Nt, including t = 0 */
    int32_t time_steps;         /* 0 sub-stepped, 1 event-driven */
    double dt;                  /* mean(diff(t)) */
    int32_t precision;          /* RTMC_PREC_* */
    int32_t block_size;         /* 0 = library default */
    int32_t particles_per_thread; /* 0 = library default (fill the device) */
    int32_t hist_mode;          /* RTMC_HIST_* */
    int32_t event_batch;        /* 0 = default (24): lanes waiting before an event step */
    int32_t n_items;
    const rtmc_batch_item* items;   /* host array of n_items */
} rtmc_batch;

/* rtmc_forward for every item of the batch in one launch.  Item i gives the
 * same histogram as rtmc_forward with item i's material, source, particles
 * and seed (the trajectories are the same; only the launch differs). */
int rtmc_forward_batched(const rtmc_batch* batch, void* stream);

/* Adjoint data of one item.  The coefficients of variation of the material
 * are c * (cv / cv[0]) with c = cv[0]; lc = CorrelationLength; Q = QP. */
typedef struct {
    const float* weight;        /* device: dL/dE per histogram bin (E's layout, raw counts) */
    double* grad;               /* device double[3]: += dL/dc, dL/dlc, dL/dQ */
    double* dhist;              /* device double[3 * hist] or NULL: += dE/dc, dE/dlc, dE/dQ
                                   per bin (hist = bins.time_stride * n_times) */
    double cv;                  /* c */
    double dlogSigma_dlc;       /* d log(Sigma) / d lc */
    const double* dlogsigma_dlc; /* device: d log sigma(theta) / d lc at theta_k = k pi / (n - 1) */
    int64_t n_dlogsigma;        /* >= 2 */
} rtmc_adjoint_item;

/* Custom backward of rtmc_forward_batched (no counterpart in the reference,
 * which has no derivatives): replays every trajectory of the batch from its
 * Philox counter and accumulates, per item, the likelihood-ratio estimate of
 * dL/d(c, lc, Q) = sum over deposits of weight[bin] * score(trajectory up to
 * the deposit), plus optionally the per-bin derivatives dE/d(c, lc, Q).
 * Acoustic media only.  `batch` must be the forward's batch (same seeds,
 * particles, materials); items' E are not touched, counters are added to. */
int rtmc_adjoint_batched(const rtmc_batch* batch, const rtmc_adjoint_item* adjoint,
                         void* stream);

/* Launch configuration chosen by rtmc_autotune_batched. */
typedef struct {
    int32_t block_size;
    int32_t particles_per_thread;
    int32_t hist_mode;
    int32_t event_batch;
    double ms;                  /* measured time of the chosen configuration on the sample */
    double ms_default;          /* measured time of the library default on the sample */
} rtmc_tuning;

/* Times candidate launch configurations (block size, particles per thread,
 * histogram privatisation, event batch) on a sample of every item
 * (sample_fraction of its particles, at least 4096) into scratch histograms,
 * and returns the fastest.  Synchronises `stream`; the items' E and counters
 * are not touched.  Histograms above 256 MiB in total are not tuned
 * (RTMC_ERR_ARG). */
int rtmc_autotune_batched(const rtmc_batch* batch, double sample_fraction,
                          rtmc_tuning* best, void* stream);

/* ------------------------------------------------------------------ */
/* Native material preparation: MaterialClass.prepareSigma
 * (MaterialClass.m:2379-2463) with calcSigmaIsotropic (:172-359) and the
 * PSDF laws (:381-710) for isotropic random media. */

/* MaterialClass.SpectralLaw */
#define RTMC_PSDF_EXP 0
#define RTMC_PSDF_POWER_LAW 1
#define RTMC_PSDF_GAUSSIAN 2
#define RTMC_PSDF_TRIANGULAR 3    /* 3-D only */
#define RTMC_PSDF_LOW_PASS 4      /* 3-D only */
#define RTMC_PSDF_VON_KARMAN 5

/* The MaterialClass properties prepareSigma reads. */
typedef struct {
    int32_t dimension;          /* 2 or 3 */
    int32_t acoustics;          /* 1 acoustic, 0 elastic */
    double v[2];                /* acoustic: v in [0]; elastic: vp, vs */
    double frequency;           /* Frequency */
    double cv[3];               /* coefficients_of_variation: acoustic kappa, rho; elastic lambda, mu, rho */
    double corr[3];             /* correlation_coefficients: acoustic [kappa-rho];
                                   elastic lambda-mu, lambda-rho, mu-rho */
    int32_t spectral_law;       /* RTMC_PSDF_* */
    int32_t pad_;
    double lc;                  /* CorrelationLength */
    double nu;                  /* SpectralParam.nu (von Karman) */
    double Q[2];                /* QP, QS; +Inf = no attenuation (acoustic: Q[0]) */
    double knot_threshold;      /* 1e-8: the reference's tables; 0: every strictly increasing knot,
                                   held beyond the last one (what the adjoint's score assumes) */
    int64_t n_theta;            /* angle grid of the pdf; 0 = 1e6 as prepareSigmaOne */
    double rho;                 /* density; > 0 (elastic): rtmc_prepare_material[s] also fill the
                                   solid/air Zoeppritz tables (n_zoe = 181); 0 = none */
    double lds_tol;             /* > 0: rtmc_prepare_material[s] also fill lds[] (rtmc_compress_table
                                   with this tol and lds_max_knots); 0 = none */
    int32_t lds_max_knots;      /* 0 = 1024 */
    int32_t pad2_;
} rtmc_material_params;

/* Host-side result of rtmc_prepare_material_host (arrays malloc'ed by the
 * library; rtmc_free_material_host releases them). */
typedef struct {
    double Sigma[4];            /* acoustic: Sigma in [0]; elastic: PP, PS, SP, SS */
    double Sigma_prime[4];      /* Sigmapr, same order */
    double mft[2];              /* meanFreeTime of P and S */
    double P2P, S2S;
    int64_t n_knots[4];         /* per table: [0] acoustic, or PP, PS, SP, SS */
    double* x[4];               /* cdf knots (GridVectors{1}) */
    double* y[4];               /* angles (Values) */
    int32_t extrap;             /* RTMC_EXTRAP_* */
    int32_t pad_;
} rtmc_material_host;

/* Replaces MaterialClass.prepareSigma for isotropic media: Sigma, the mean
 * free times, P2P / S2S and the inverse-CDF tables, on the host. */
int rtmc_prepare_material_host(const rtmc_material_params* params, rtmc_material_host* out);
void rtmc_free_material_host(rtmc_material_host* out);

/* The same, uploaded: fills `out` (device tables) for rtmc_forward and the
 * batch entry points; *handle owns the device memory until
 * rtmc_free_material.  An elastic material with params->rho > 0 also gets
 * the Zoeppritz tables of elastic reflective boundaries
 * (rtmc_zoeppritz_air on 181 angles); otherwise out->n_zoe = 0. */
int rtmc_prepare_material(const rtmc_material_params* params, rtmc_material* out, void** handle);
int rtmc_free_material(void* handle);

/* No counterpart in the reference (it prepares one material per call): n
 * materials prepared at once for the batch entry points, one material per
 * host thread, all tables in one device block owned by *handle
 * (rtmc_free_material).  Errors name the first failing material. */
int rtmc_prepare_materials(const rtmc_material_params* params, int32_t n, rtmc_material* out,
                           void** handle);

/* Replaces MaterialClass.Zoeppritz (MaterialClass.m:1786-1830) for its one
 * case, a solid under air (vp 343, vs 0, rho 1; ZoeppritzFluid.m:1-245):
 * at the n incidence angles deg = linspace(0, 90, n) (the reference uses 181),
 * the reflected energy fractions E_Rpp (P incidence) and E_Rsvsv (SV
 * incidence) of a solid with velocities vp, vs and density rho.  Host arrays
 * of n values each. */
int rtmc_zoeppritz_air(double vp, double vs, double rho, int64_t n, double* deg, double* E_Rpp,
                       double* E_Rsvsv);

/* No counterpart in the reference: compresses one inverse-CDF table
 * (griddedInterpolant(x, y) with linear extrapolation, MaterialClass.m:2459-2462;
 * host arrays) into at most max_knots of its own knots for LDS.  Knots 0, 1,
 * n-2 and n-1 are always kept; between them, a run of knots is dropped while
 * the chord across it stays within tol of every dropped knot (both functions
 * are piecewise linear, so then |theta(q) - table(q)| <= tol for every q).
 * When more than max_knots are needed, tol is doubled until they fit;
 * *tol_used is the bound achieved.  out: host array of max_knots. */
int rtmc_compress_table(const double* x, const double* y, int64_t n, double tol, int32_t max_knots,
                        rtmc_lds_knot* out, int32_t* m, double* tol_used);

/* No counterpart in the reference: fills mats[i].lds[t] for every inverse-CDF
 * table of n prepared materials (device tables, read back on `stream`), by
 * rtmc_compress_table(tol, max_knots), one material per host thread; the
 * knots live in one device block owned by *handle (rtmc_free_material).  The
 * tracing kernel then holds them in LDS and draws every scattering angle
 * from them, when they fit beside the launch's histogram (else from invcdf;
 * RTMC_LDS_TABLE=0 forces invcdf). */
int rtmc_attach_lds_tables(rtmc_material* mats, int32_t n, double tol, int32_t max_knots, void** handle,
                           void* stream);

/* No counterpart in the reference: the material derivatives the adjoint
 * (rtmc_adjoint_batched) needs, by central differences of calcSigmaIsotropic
 * (MaterialClass.m:204-235) at lc (1 +- rel), acoustic media:
 * dlogsigma_dlc[i * n_theta + k] = d log sigma(theta_k) / d lc on
 * theta = linspace(0, pi, n_theta) (0 where sigma is not positive), and
 * dlogSigma_dlc[i] = d log Sigma / d lc. */
int rtmc_lc_derivatives(const rtmc_material_params* params, int32_t n, int64_t n_theta, double rel,
                        double* dlogsigma_dlc, double* dlogSigma_dlc);

/* Measurement aid, no counterpart in the reference: throughput of no-return
 * global atomics (requests/s) on the current device, the ceiling of the
 * fused kernel's deposits when the histogram is in global memory.  buf:
 * device uint32[words] (contents are changed).  kind 0: uint32 at random
 * words; 1: uint64 at random aligned pairs; 2 / 3: runs of 2 / 8 consecutive
 * words from one lane; 4: as 1, returning the old values.  Synchronises `stream`. */
int rtmc_probe_atomics(uint32_t* buf, int64_t words, int32_t kind, int64_t iters,
                       double* requests_per_s, void* stream);

/* Measurement aid, no counterpart in the reference: when on, every forward
 * or adjoint call on this thread records HIP events around its tracing
 * kernel on the call's stream (at most 1024 launches are kept).
 * rtmc_trace_times waits for them and returns their durations in ms (at
 * most `max`, count in *n), then forgets them. */
int rtmc_set_trace_timing(int32_t on);
int rtmc_trace_times(double* ms, int32_t max, int32_t* n);

#ifdef __cplusplus
}
#endif

#endif /* RTMC_H */

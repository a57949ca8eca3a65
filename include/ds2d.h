/*
 * ds2d.h — C-ABI of the MI355X-native DS2D time-domain hot path (libds2d.so).
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point
 * returns 0 on success and a negative DS2D_E* code on failure; the message is
 * available from ds2d_last_error() (thread-local).  One handle may be used by
 * one host thread at a time; all device work of a handle runs on its own HIP
 * stream.  Host arrays are owned by the caller; device buffers are owned by
 * the handle (allocated in ds2d_create / on first use, freed in ds2d_destroy).
 *
 * Layouts follow the reference: models and images are [ix][iz] with z
 * contiguous (allnx*allnz floats, x-major); records are [trace][it].
 *
 * Reference interfaces replaced (upstream DS2D tree, file:line):
 *   ds2d_create            FD2DGPU_ELASTIC::FD2DGPU_ELASTIC  FD2DGPU.cuh:130-131, FD2DGPU.cu:1350
 *   ds2d_destroy           FD2DGPU_ELASTIC::~FD2DGPU_ELASTIC FD2DGPU.cu:1636
 *   ds2d_set_model         GPUbufferVPVS                     FD2DGPU.cu:1311-1315 (main.cpp:940)
 *   ds2d_set_image         GPUbufferM                        FD2DGPU.cu:1318-1322 (main.cpp:954)
 *   ds2d_set_shot          record_copytoGPU + bufferVpVsHtoD + public fields
 *                                                            FD2DGPU.cu:3357-3383, 1336-1348;
 *                                                            main.cpp:994-1025
 *   ds2d_shot_gradient     ELSRTM_ROT_EXP_IMAGE_SINGLESHOT   FD2DGPU.cuh:429-430, FD2DGPU.cu:4184-4879
 *   ds2d_take_gradients    imagebuffer_resettozero           FD2DGPU.cu:3478-3505 (main.cpp:1035)
 *   ds2d_forward           elastic_rotation_addS_for + born_modeling(1,0) + GPUseisrecord
 *                                                            FD2DGPU.cu:2046-2049, 1834-1839, 355-365
 *   ds2d_batch_gradient    the per-worker shot loop of lsrtm_all_iteration with the
 *                          MPI shot dispatch replaced by one batched launch per step
 *                                                            main.cpp:956-1036
 *   ds2d_batch_forward     FD_ELASTIC forward modelling (FD2DGPU.cu:2241, dead) batched
 *   ds2d_acoustic_ref_forward  FD() calculateP/calculateV/addS/recordD FD2DGPU.cu:523-583
 *   ds2d_born_forward / ds2d_born_adjoint / ds2d_lsrtm_cg
 *                          plain-L2 counterpart of the Born/adjoint pair inside
 *                          ELSRTM_ROT_EXP_IMAGE_SINGLESHOT and of sjcgsolver
 *                          (main.cpp:1224-1320); the reference has no exact adjoint
 *   ds2d_laplace_filter    laplace_filter_2d(nz,nx,order,dx,dat) laplace.cpp:82-106
 *   ds2d_laplace_band      laplace_filter_2d(nz,nx,dat)          laplace.cpp:3-57
 *   ds2d_ricker            ricker1                               main.cpp:461-473
 *   ds2d_sjcg_update       sjcgbeta + sjcglength + sjcgsolver    main.cpp:1232-1303
 *   ds2d_rot_forward       rot_exp_calculate_v/_p + elastic_rotation_addS_for
 *                                                            FD2DGPU.cu:3387-3424, 2046-2049
 *   ds2d_rot_rtm           ROT_EXP_IMAGE_SINGLESHOT (commented out) FD2DGPU.cu:3505-3585
 *   ds2d_decomp_rtm        FD_ELASTIC with VECTORIMG / POYNTING  FD2DGPU.cu:2241-2517
 */
#ifndef DS2D_H
#define DS2D_H

#ifdef __cplusplus
extern "C" {
#endif

#define DS2D_OK 0
#define DS2D_EINVAL (-1)   /* bad argument / inconsistent sizes            */
#define DS2D_EDEVICE (-2)  /* HIP runtime error or no usable gfx950 device */
#define DS2D_ENOMEM (-3)   /* device allocation failed                     */
#define DS2D_ESTATE (-4)   /* call out of order (no model, no shot, ...)    */
#define DS2D_ESHOT (-5)    /* some shots of a batch failed (non-finite); the
                              others completed: see ds2d_shot_status        */

typedef struct ds2d_ctx ds2d_t;

/* Constructor arguments of FD2DGPU_ELASTIC (main.cpp:939) plus the options
 * the reference hard-codes.  nxpml/nzpml describe the largest window; a batch
 * may hold narrower windows. */
typedef struct {
    float dx, dz, dt;
    int nxpml, nzpml;       /* padded window (largest)                         */
    int allnx, allnz;       /* global model                                    */
    int scale;              /* trace decimation of the surface recorder (1)    */
    int pml;                /* frame width (PML or random boundary)            */
    int nt;                 /* time steps                                      */
    int nop;                /* half stencil length; must be 4                  */
    int ntr;                /* traces per shot (largest)                       */
    int physics;            /* 0 elastic, 1 acoustic                           */
    float pml_rr;           /* PML reflection coefficient (elastic 1e-3)       */
    float rho_const;        /* > 0: constant density (reference: 2000)         */
    float rb_amp;           /* random boundary relative amplitude (0.3)        */
    unsigned long long rb_seed; /* random boundary seed                        */
    int bg_mode;            /* plain Born/adjoint background: 0 auto, 1 keep the
                               per-step increment resident in HBM, 2 rebuild it
                               by reversal through the random boundary          */
    float hbm_budget_gb;    /* auto mode: largest history kept resident (64)   */
} ds2d_geom;

/* One shot / aperture window (main.cpp:1011-1025 public fields). */
typedef struct {
    int ngx_min;            /* window start column in the global model          */
    int nx;                 /* window width (physical cells)                    */
    int isx, isz;           /* source cell, padded-window coordinates           */
    int igz;                /* receiver row, padded coordinates (pml + gy)      */
    int ntr;                /* number of traces                                 */
    const int* rec_ix;      /* [ntr] receiver columns, padded coordinates       */
    int shot_id;            /* seeds the random boundary of this shot           */
    int boundary;           /* 0 PML, 1 random boundary                         */
} ds2d_shot;

/* Timing of the last run (HIP events on the handle's stream). */
typedef struct {
    double elapsed_ms;         /* whole run                                       */
    double stencil_ms;         /* time-stepping loop only                         */
    long long cell_updates;    /* sum over wavefields of nxpml*nzpml*nt            */
    long long stencil_launches;/* launches of the dominant stencil kernel          */
    double alg_bytes;          /* algorithmic HBM bytes of one such launch         */
    double kernel_avg_ms;      /* its average launch duration (sampled events)     */
    long long kernel_samples;  /* launches that were bracketed by events           */
    double aux_alg_bytes;      /* the other kernel of a two-kernel step, per launch */
    double aux_kernel_avg_ms;  /* (0 when the step is one kernel)                  */
    double blocked_bytes;      /* several time steps per launch: the bytes one launch
                                  itself must move (alg_bytes counts each step's
                                  cell-updates at the one-step figure); else 0     */
} ds2d_stats;

const char* ds2d_last_error(void);
int ds2d_device_info(char* name, int len, int* n_cu, double* hbm_gb);

int ds2d_create(const ds2d_geom* g, const float* wavelet, ds2d_t** out);
int ds2d_destroy(ds2d_t* h);

int ds2d_set_model(ds2d_t* h, const float* vp_all, const float* vs_all, const float* rho_all);
int ds2d_set_image(ds2d_t* h, const float* pp, const float* ps);
int ds2d_set_shot(ds2d_t* h, const ds2d_shot* s, const float* rec_z, const float* rec_x);

/* Reference per-shot LSRTM gradient (compat mode: Born + 2nd-order Born,
 * NCC residual, three-term imaging).  Accumulates into the handle's images. */
int ds2d_shot_gradient(ds2d_t* h, int iter, double* misfit);
/* Same for many shots in one batched launch per step (obs_* concatenated).
 * A shot whose NCC residual or adjoint images hold a NaN or an infinity is
 * left out of the misfit and of the accumulated images, the other shots
 * complete, and the call returns DS2D_ESHOT (no reference counterpart: the
 * reference has no per-shot failure handling). */
int ds2d_batch_gradient(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* obs_z,
                        const float* obs_x, int iter, double* misfit);
/* Per-shot status of the last ds2d_batch_gradient / ds2d_shot_gradient
 * (nshots must match it): 0 = included, 1 = left out (non-finite). */
int ds2d_shot_status(ds2d_t* h, int nshots, int* status);
int ds2d_take_gradients(ds2d_t* h, float* pp, float* ps, float* illum, float* b_pp, float* b_ps,
                        float* b2_pp, float* b2_ps);
/* Precision of the reference-compat gradient's two source-side wavefields
 * (the background and the first-order Born field, whose velocities drive
 * add_born_source): 1 = FP64 (default), 0 = FP32 as the reference.  Every
 * other wavefield, the records and the images are FP32 either way.  No
 * reference counterpart (the reference computes all fields in FP32). */
int ds2d_set_precision(ds2d_t* h, int source_fp64);
/* Reference Born modelling alone, batched: order 1 returns the first-order
 * scatter records Lm (born_modeling mode 1 + record_scatter, FD2DGPU.cu:1820-1830,
 * 3460-3467), order 2 returns Lm + Lm2 (modes 0/3/4/5 + scatter_add,
 * FD2DGPU.cu:1834-1860, 3147-3152) -- the records the gradient's NCC consumes.
 * The image is the handle's (ds2d_set_image); records concatenated in shot order. */
int ds2d_born(ds2d_t* h, int nshots, const ds2d_shot* shots, int order, float* rec_z, float* rec_x);
/* ELSRTM_CALCULATE_ALPHA (FD2DGPU.cu:4885-5035): with m the handle's image and
 * g the given gradient images, r = Lm - d and q = Lg (first-order Born);
 * *num += <q, r>, *den += <q, q> in FP64 over all shots (fenzi, fenmu).  The
 * handle's image is left as it was. */
int ds2d_calculate_alpha(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* g_pp, const float* g_ps,
                         const float* obs_z, const float* obs_x, double* num, double* den);

/* Elastic forward modelling of the current shot: vz, vx at (rec_ix, igz). */
int ds2d_forward(ds2d_t* h, float* rec_z, float* rec_x);
/* Batched forward modelling of many uneven windows.  Records concatenated in
 * shot order, each [ntr_s][nt].  mode: 0 = reference split-field kernels,
 * 1 = two fast kernels per step (unsplit interior), 2 = one fused kernel per
 * step (or what ds2d_autotune_forward picked). */
int ds2d_batch_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, float* rec_z, float* rec_x,
                       int mode, ds2d_stats* stats);
/* Tile of the fused kernel: x-chunk width (0 = default 64), workgroup size
 * (256, 512, 704 or 768 lanes, one z row each, owning threads-16 rows; 0 =
 * default 256) and prefetch depth (columns of operands in flight, 2 or 4;
 * 0 = 2; two steps per launch always run with 2); and an autotuner that times
 * nsteps per candidate on this batch -- the fused kernel at several tiles,
 * one or two time steps per launch, and the two-kernel step (reported as
 * width 0) -- and keeps the fastest for mode 2.  No reference counterpart
 * (new capability). */
int ds2d_set_chunk(ds2d_t* h, int width, int threads, int depth);
/* Time steps per launch of the fused elastic forward (mode 2): 1, or 2 =
 * temporal blocking (steps it and it+1 in one pass over HBM; workgroups of
 * 512, 704 or 768 lanes own threads-32 rows).  Set by the autotuner too.
 * No reference counterpart (new capability). */
int ds2d_set_time_block(ds2d_t* h, int steps);
int ds2d_get_time_block(ds2d_t* h, int* steps);
/* Chunk list of two steps per launch: on (1, the default) the library plans
 * the chunks of the batch for balance across the eight XCDs and their CUs
 * (each XCD gets the same number of chunks, a multiple of its CU count,
 * windows split into equal-width chunks, widest first) whenever a simulated
 * list schedule says that finishes sooner than fixed-width chunks of the set
 * width; off (0) always uses the fixed width.  Results are identical either
 * way.  ds2d_get_chunk_plan reports whether the last two-step launch list
 * was the balanced plan (1) or fixed width (0).  No reference counterpart
 * (new capability). */
int ds2d_set_chunk_plan(ds2d_t* h, int on);
int ds2d_get_chunk_plan(ds2d_t* h, int* used);
/* Forward pass of the plain-L2 elastic Born pair (ds2d_eborn_forward and the
 * Born passes of ds2d_elsrtm_cg / ds2d_elsrtm_learned): on (1, the default)
 * one fused launch per time step advances the background and the scattered
 * field (velocity, stress, Born source, recorder) with the arithmetic of the
 * two per-step kernels (equal to FP32 rounding); off (0) launches those two
 * kernels.
 * width: x-chunk width of the fused step (0 = 64); threads: 256 or 512 lanes
 * (0 = 256).  No reference counterpart (new capability). */
int ds2d_set_fused_born(ds2d_t* h, int on, int width, int threads);
/* Rotation-scheme forward modelling (ds2d_rot_forward): on (1, the default)
 * one fused launch per time step updates the velocity parts and theta /
 * omega (unsplit outside the damping strips); off (0) launches the per-cell
 * k_rot_v and k_rot_p of the reference layout.  No reference counterpart
 * (new capability). */
int ds2d_set_fused_rotation(ds2d_t* h, int on);
/* Capture each time loop into one hipGraph (1) or launch step by step (0). */
int ds2d_set_graph(ds2d_t* h, int on);
int ds2d_autotune_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, int nsteps, int* best_width,
                          int* best_threads, int* best_depth, double* best_ms);

/* Reference acoustic step (split p, mirrored stagger): pressure at depth pml. */
int ds2d_acoustic_ref_forward(ds2d_t* h, int sxnum, int ngx_min, float* rec_p);

/* Plain-L2 acoustic Born operator L (refl -> records) and its exact adjoint
 * L^T (residual -> gradient, allnx*allnz, overwritten), batched over shots.
 * Random boundary, no damping.  rec/res concatenated [shot][trace][it]. */
int ds2d_born_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* refl, float* rec_bg,
                      float* rec_born, ds2d_stats* stats);
int ds2d_born_adjoint(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* res, float* grad,
                      ds2d_stats* stats);

/* Conjugate-gradient LSRTM (CGLS on the plain Born operator).  image in/out
 * (allnx*allnz); resid[niter+1] receives |Lm-d|/|d| before each iteration and
 * after the last. */
int ds2d_lsrtm_cg(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* d_obs, int niter,
                  float* image, double* resid, ds2d_stats* stats);

/* Plain-L2 elastic Born operator L: (PP, PS) images -> scattered (vz, vx)
 * records, and its exact adjoint (records -> PP, PS gradients, overwritten),
 * batched over shots; physics 0, random boundary, no damping.  The Born
 * source is add_born_source (born.cu:351-405, direction 1) on the physical
 * window and the imaging products are image_gradient's (FD2DGPU.cu:3336-3339);
 * the propagation replaces the reference's born_modeling / GPUcalculate_*_back
 * schedule (FD2DGPU.cu:1817-1897, 591-895).  Images allnx*allnz; records
 * [shot][trace][it]. */
int ds2d_eborn_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* img_pp, const float* img_ps,
                       float* rec_z, float* rec_x, ds2d_stats* stats);
int ds2d_eborn_adjoint(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* res_z, const float* res_x,
                       float* g_pp, float* g_ps, ds2d_stats* stats);
/* ELSRTM_CALCULATE_ALPHA (FD2DGPU.cu:4885-5035, uncalled in the reference) on
 * the plain elastic pair: r = L m - d, q = L g, *num = <q, r>, *den = <q, q>
 * (FP64 over all shots; the step of m - alpha g is alpha = num / den). */
int ds2d_eborn_line_search(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* img_pp, const float* img_ps,
                           const float* g_pp, const float* g_ps, const float* d_z, const float* d_x, double* num,
                           double* den);
/* CGLS PP/PS LSRTM on the elastic Born pair; images in/out, resid[niter+1] as
 * for ds2d_lsrtm_cg. */
int ds2d_elsrtm_cg(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* d_z, const float* d_x, int niter,
                   float* img_pp, float* img_ps, double* resid, ds2d_stats* stats);

/* ---------------------------------------------------------------------------
 * Learned LSRTM preconditioner (no reference counterpart: new capability).
 * A U-Net (levels, base channels; 3x3 convs + ReLU, 2x2 average pooling,
 * nearest upsampling with skip concatenation, 1x1 head) mapping an RTM image
 * to a multi-iteration LSRTM image.  Weights and activations are BF16 on
 * the MFMA units with FP32 accumulation; master weights, gradients and Adam
 * moments are FP32.  Images are [n][h][w] floats (an [ix][iz] image is
 * h = nx, w = nz); h and w must be multiples of 2^(levels-1).
 * Parameters are flattened layer by layer (enc0a, enc0b, ..., enc{L-1}b,
 * dec{L-2}a, dec{L-2}b, ..., dec0b, head), each weight ([tap][cout][cin],
 * tap = (dx+1)*3 + (dz+1)) followed by its bias.
 * ------------------------------------------------------------------------- */
typedef struct ds2d_unet ds2d_unet_t;
typedef struct {
    int levels;              /* 4 */
    int base;                /* base channels, 32 (multiple of 32)            */
    int max_batch, max_h, max_w;
    unsigned long long seed; /* He-uniform initialisation, zero biases         */
} ds2d_unet_cfg;
typedef struct {
    double elapsed_ms;       /* whole call                                     */
    double ms_per_step;      /* training: per Adam step; inference: per call   */
    double flops_per_step;   /* algorithmic conv FLOPs (2 per MAC)             */
    int steps;
} ds2d_unet_stats;

int ds2d_unet_create(const ds2d_unet_cfg* cfg, ds2d_unet_t** out);
int ds2d_unet_destroy(ds2d_unet_t* u);
int ds2d_unet_num_params(const ds2d_unet_t* u, long long* n);
int ds2d_unet_get_params(ds2d_unet_t* u, float* theta);
int ds2d_unet_set_params(ds2d_unet_t* u, const float* theta); /* also resets Adam */
int ds2d_unet_forward(ds2d_unet_t* u, int n, int h, int w, const float* x, float* y, ds2d_unet_stats* stats);
/* Mean-squared-error loss of forward(x) against t and its gradient. */
int ds2d_unet_loss_grad(ds2d_unet_t* u, int n, int h, int w, const float* x, const float* t, double* loss,
                        float* grad);
/* Keep a training set resident in HBM, then run Adam steps on random batches
 * (seeded, without replacement within an epoch); loss_hist[nsteps]. */
int ds2d_unet_set_data(ds2d_unet_t* u, int nsamples, int h, int w, const float* x, const float* t);
int ds2d_unet_train(ds2d_unet_t* u, int nsteps, int batch, float lr, unsigned long long seed, float* loss_hist,
                    ds2d_unet_stats* stats);
/* One BF16 MFMA 3x3 convolution kernel on host FP32 data (inputs rounded to
 * BF16 on the device, FP32 outputs).  op 0: y = conv(concat(x1, up^up(x2)))
 * [+ bias] [ReLU] -> out [npix][cout]; op 1: data gradient of that conv,
 * x1 = dy [npix][cout] -> out [npix][c1 + c2]; op 2: weight gradient, dy as
 * extra input -> out [9][cout][c1 + c2].  kernel_ms: average duration of the
 * conv kernel over `repeat` launches (HIP events). */
int ds2d_conv3x3_bf16(int op, int n, int h, int w, int c1, int c2, int up, int cout, const float* x1,
                      const float* x2, const float* dy, const float* wt, const float* bias, int relu, int repeat,
                      float* out, double* kernel_ms);

/* Learned LSRTM: flexible preconditioned CG on the plain Born operator in which
 * the direction is the U-Net applied to the (RMS-normalised) gradient.  A step
 * whose exact line search lowers the data residual by less than 1% (or a
 * non-descent direction) is discarded and the solver continues as plain CGLS
 * from the current image.  resid[niter+1] as ds2d_lsrtm_cg; *learned_steps
 * receives the number of accepted learned steps. */
int ds2d_lsrtm_learned(ds2d_t* h, ds2d_unet_t* u, int nshots, const ds2d_shot* shots, const float* d_obs,
                       int niter, float* image, double* resid, int* learned_steps, ds2d_stats* stats);
/* The same on the plain elastic PP/PS Born pair (ds2d_eborn_*): the U-Net
 * preconditions the PP and the PS gradient images one at a time. */
int ds2d_elsrtm_learned(ds2d_t* h, ds2d_unet_t* u, int nshots, const ds2d_shot* shots, const float* d_z,
                        const float* d_x, int niter, float* img_pp, float* img_ps, double* resid, int* learned_steps,
                        ds2d_stats* stats);
/* Launch shots in groups of at most n per time loop (0 = all shots in one
 * batched launch, the default; 1 = one shot after another as the reference's
 * per-rank shot loop does).  Applies to the plain Born/adjoint/CG entries. */
int ds2d_set_batch(ds2d_t* h, int shots_per_launch);
/* Resident-history Born and adjoint passes as one fused launch per step
 * (x-marching workgroups): -1 = auto (default: the per-cell kernels and the
 * fused step at several x widths are timed once per batch geometry), 0 =
 * never (per-cell kernels), 1 = always (x width 64), k in [8, 1024] = always,
 * with x width k. */
int ds2d_set_fused(ds2d_t* h, int mode);
/* Stop rule of the CG solvers (ds2d_lsrtm_cg, ds2d_lsrtm_learned,
 * ds2d_elsrtm_cg, ds2d_elsrtm_learned): return after the first iteration
 * whose relative data residual |d - Lm|/|d| is at most target (the rest of
 * resid repeats it); 0 = run all niter iterations (default).  No reference
 * counterpart (the reference runs a fixed iteration count). */
int ds2d_set_stop(ds2d_t* h, double target);
/* The CG solvers print "ds2d: iteration k, relative data residual r" to
 * stderr every `every` iterations and at the last one (0 = silent, the
 * default), as the reference prints its misfit per iteration. */
int ds2d_set_log(ds2d_t* h, int every);

/* Host-side nonlinear CG update of the reference (sjcgsolver, main.cpp:1232-1303):
 * cg = -g1 + beta cg, beta = max(0, min(HS, DY)) (0 at iter 0); g0 <- g1;
 * alpha = 1e-4 at iter 0, else 0.1 max|m| / max|cg|; m += alpha cg. */
int ds2d_sjcg_update(int n, float* m, float* cg, const float* g1, float* g0, int iter, float* alpha,
                     float* beta);

int ds2d_laplace_filter(int nz, int nx, int order, float dx, float* dat);
int ds2d_laplace_band(int nz, int nx, float* dat);
/* ds2d_laplace_band on device buffers that allocates nothing: no counterpart
 * in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  `in` and `out` are device
 * pointers of nx*nz floats (they may be equal); the work is queued on `stream`
 * (a hipStream_t, or null for the default stream) and the call returns
 * without waiting.  The two intermediate images live in a library-owned
 * device array of 2 x 4194304 floats per GPU (held until the process ends):
 * images above 4194304 cells are rejected with DS2D_EINVAL.  Calls that
 * overlap on one GPU share that array and race. */
int ds2d_laplace_band_noalloc(int nz, int nx, const float* in, float* out, void* stream);
int ds2d_ricker(int nt, float f0, float dt, float* out);

/* P/S-decoupled (theta/omega rotation) elastic modelling of a batch: per step
 * theta += s[it] at the source (elastic_rotation_addS_for, FD2DGPU.cu:2046-2049,
 * theta aliased to txx at 1407), GPUcalculate_rotation_v then _p
 * (FD2DGPU.cu:3005-3080; undamped forms 2545-2609), then vz, vx at
 * (rec_ix, igz).  PML or random boundary per shot.  The medium is vp and vs
 * (the scheme carries vp^2 and vs^2, no density).  Records concatenated in
 * shot order, each [ntr_s][nt].  stats may be null. */
int ds2d_rot_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, float* rec_z, float* rec_x,
                     ds2d_stats* stats);
/* P/S-decomposed RTM with the rotation scheme (ROT_EXP_IMAGE_SINGLESHOT,
 * FD2DGPU.cu:3505-3585): the source wavefield is rebuilt by reversal through
 * the random boundary (every shot needs boundary = 1); the receiver side
 * adds rec_z / rec_x to vz / vx at the receivers (FD2DGPU.cu:3306-3327); at
 * every step PP += (vpx vpx' + vpz vpz') / (vpx^2 + vpz^2 + 1e-6) and PS the
 * same with the receiver's S parts (doimage, FD2DGPU.cu:3218-3233), stacked
 * over shots into allnx*allnz images (stack_image, FD2DGPU.cu:3082-3091).
 * Deviation: the reversal re-injects each source sample before the velocity
 * update, which makes it exact (the reference's order injects it twice).
 * Both wavefields run in FP64 (the image divides by the source's own energy,
 * which magnifies FP32 rounding where the source field is small). */
int ds2d_rot_rtm(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* rec_z, const float* rec_x,
                 float* img_pp, float* img_ps);
/* Elastic RTM with decomposed imaging (FD_ELASTIC, FD2DGPU.cu:2241-2517):
 * source side propagated forward with the live elastic step, its imaging
 * quantity kept per step in device memory (the reference stages it on disk);
 * receiver side propagated from rest with txx = tzz = rec_z[tr][it] assigned
 * at the receivers (GPUapplysource_backward, FD2DGPU.cu:2128-2134) for
 * it = nt-1 .. 0, imaged at every step but the first.  Per step, on the
 * physical window: Helmholtz VP/VS (FD2DGPU.cu:1963-1986), vector VP/VS
 * (1988-2017), and by kind
 *   0  PP, vector P split up/down by the Poynting flux (2223-2236),
 *      (fu.bd + fd.bu) / (fu.fu + fd.fd + 1e-6)  (doimage_poynting, 2164-2196)
 *   1  PP, 2 PS, 3 SS vector images: f.b / (f.f + 1e-6)   (doimage, 2141-2162)
 *   4  PP, 5 PS, 6 SS scalar images: f b / (f^2 + 1e-6)
 * (PS: the source's P against the receiver's S).  Stacked over shots into
 * one allnx*allnz image.  Device memory: nt x window cells x (4, 2 or 1)
 * floats for the source side. */
int ds2d_decomp_rtm(ds2d_t* h, int nshots, const ds2d_shot* shots, int kind, const float* rec_z, float* image);

#ifdef __cplusplus
}
#endif
#endif /* DS2D_H */

/*
 * gigalens_amd -- C ABI of the MI355X (gfx950) lens-model hot path.
 *
 * The ABI is what a binding of the reference's LensSimulator / prob-model
 * path would call.  Every tensor argument is caller-allocated DEVICE memory
 * (row-major float32 unless stated), every call is stream-ordered and
 * asynchronous.  Only the create / destroy calls allocate or free device memory
 * and synchronise the host: gl_plan_create, gl_plan_destroy, gl_batch_create,
 * gl_batch_create_ws, gl_batch_create_ex, gl_batch_destroy, gl_cnn_create, gl_cnn_destroy.
 * gl_cnn_set_params clears the AdamW moments with hipMemsetAsync and must not be recorded
 * into a HIP graph.
 * Errors are reported as a non-zero return code; gl_last_error() returns a
 * thread-local message.
 * A plan, a batch and a network own scratch memory (the staged pipeline's
 * workspace, the fused kernel's launch-order buffer, the network's activations)
 * that their calls share.  The library orders the calls on one object across
 * streams: a call on another stream than the previous call makes its stream wait
 * for the work queued on the previous one (an event wait; calls on one stream add
 * nothing), so concurrent streams on one object are serialised, not raced; a
 * stream under graph capture is not ordered this way (order the work before the
 * capture).  Use one plan per stream for concurrent work.
 *
 * Parameter vectors use the reference's flatten order (tf.nest order of a
 * prior sample: component lists in order, dict keys sorted), i.e. the
 * column order of z in src/gigalens/tf/model.py:40-51.
 *
 * Citations are file:line in the reference repository.
 */
#ifndef GIGALENS_AMD_H
#define GIGALENS_AMD_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GL_OK 0
#define GL_ERR_ARG 1
#define GL_ERR_HIP 2
#define GL_ERR_UNSUPPORTED 3

#define GL_MAX_COMPONENTS 8
#define GL_MAX_COMPONENT_PARAMS 10
#define GL_MAX_PARAMS 128
#define GL_MAX_EPL_ITER 100

/* Profile kinds (src/gigalens/tf/profiles/{mass,light}/...). */
enum gl_profile_kind {
    GL_EPL = 1,            /* params: theta_E, gamma, e1, e2, center_x, center_y  (epl.py:17) */
    GL_SHEAR = 2,          /* params: gamma1, gamma2                              (shear.py:10) */
    GL_SIS = 3,            /* params: theta_E, center_x, center_y                 (sis.py:11) */
    GL_SIE = 4,            /* params: theta_E, e1, e2, center_x, center_y         (sie.py:12) */
    GL_TNFW = 5,           /* params: Rs, alpha_Rs, r_trunc, center_x, center_y   (tnfw.py:12) */
    GL_SERSIC = 10,        /* params: R_sersic, n_sersic, center_x, center_y, Ie   (sersic.py:23) */
    GL_SERSIC_ELLIPSE = 11,/* params: R_sersic, n_sersic, e1, e2, center_x, center_y, Ie (sersic.py:64) */
    GL_SHAPELETS = 12,     /* params: beta, center_x, center_y; amplitudes solved by least squares
                              (use_lstsq) or sampled at columns amp_col0.. (shapelets.py:17,53-75);
                              HBM-staged pipeline only */
    GL_CORE_SERSIC = 13    /* params: R_sersic, n_sersic, Rb, alpha, gamma, e1, e2, center_x, center_y, Ie
                              (sersic.py:75-88) */
};

/* Roles: which list of PhysicalModel the component belongs to (model.py:21-29). */
enum gl_role { GL_ROLE_LENS = 0, GL_ROLE_LENS_LIGHT = 1, GL_ROLE_SOURCE = 2 };

/* Prior families of the reference default prior (tests/conftest.py:20-73). */
enum gl_prior_kind {
    GL_PRIOR_NORMAL = 1,        /* args: loc, scale          bijector Identity */
    GL_PRIOR_LOGNORMAL = 2,     /* args: loc, scale          bijector Exp */
    GL_PRIOR_UNIFORM = 3,       /* args: low, high           bijector Sigmoid(low, high) */
    GL_PRIOR_TRUNCNORMAL = 4    /* args: loc, scale, low, high  bijector Sigmoid(low, high) */
};

enum gl_likelihood {
    GL_LIK_FORWARD = 0,   /* sigma from the simulated image (tf/model.py:77) */
    GL_LIK_BACKWARD = 1   /* sigma from the observed image  (tf/model.py:113-115) */
};

typedef struct gl_component_desc {
    int kind;                                   /* gl_profile_kind */
    int role;                                   /* gl_role */
    int param_col[GL_MAX_COMPONENT_PARAMS];     /* column of x for each profile parameter, -1: constant */
    float param_const[GL_MAX_COMPONENT_PARAMS]; /* value used when param_col is -1 */
    int niter;                                  /* EPL: maximum series length (epl.py:19) */
    int n_max;                                  /* SHAPELETS: n_max */
    int interpolate;                            /* SHAPELETS: table mode (shapelets.py:20,55-65) */
    int use_lstsq;                              /* light: unit amplitude, solved by least squares */
    int amp_col0;                               /* SHAPELETS (not lstsq): column of amp00, contiguous */
} gl_component_desc;

typedef struct gl_plan_desc {
    /* camera: src/gigalens/simulator.py:11-28, tf/simulator.py:17-55 */
    int num_pix;
    int supersample;
    float transform_sub[4];   /* sub-pixel pix->angle matrix (T / ss), row-major */
    float origin[2];          /* angle at sub-pixel (0,0): ra_at_xy_0, dec_at_xy_0 */
    float conversion_factor;  /* det(T) (tf/simulator.py:24-26) */
    int kernel_size;          /* k of the sub-grid PSF (0: no PSF) */
    const float* kernel;      /* HOST k*k sub-grid PSF, NOT flipped (true convolution) */
    /* model */
    int n_components;
    const gl_component_desc* components; /* HOST; lens comps first, then lens light, then source */
    /* parameters / prior */
    int n_params;             /* d */
    const int* prior_kind;    /* HOST d entries, gl_prior_kind */
    const float* prior_args;  /* HOST d*4 */
    /* data */
    const float* observed;    /* HOST N*N, may be NULL for simulate-only plans */
    float background_rms;
    float exp_time;
    int likelihood;           /* gl_likelihood */
    int fldj_at_x;            /* 1: evaluate fldj at constrained x (tf/model.py:160 quirk) */
    int epl_batch_series;     /* 1: reference's batch-coupled EPL series length (epl.py:41) */
    float psf_rank_tol;       /* separable PSF: keep singular values >= tol * s_max of the folded
                                 PSF; 0 selects 1e-6, a negative value forces the direct 2-D path.
                                 Staged plans (and lens batches of them) use the same rule for
                                 their channel convolution and adjoint convolution */
    int max_batch;            /* staged plans: particles per internal chunk (the workspace is
                                 allocated for this many at plan creation), at most what a 64 GiB
                                 workspace holds; 0 selects as many as fit 32 GiB, at most 1024 */
    const float* noise_map;   /* HOST N*N sigma per pixel; NULL: sqrt(bkg^2 + clip(obs,0)/t_exp) */
    int n_observed;           /* observed images stacked in `observed` (and `noise_map`), each
                                 N*N; 0 selects 1.  Particles pick one with the lens index of
                                 gl_log_prob_lenses / gl_map_adam_step_lenses */
    int staged;               /* 1: always use the HBM-staged pipeline (otherwise it is chosen when
                                 the LDS-resident kernel cannot hold the image, has no compiled
                                 kernel for the profile combination, or lights are linear) */
    int conv_rows;            /* staged plans: output rows per lane of the channel convolution
                                 (2, 4 or 8; 0 selects 4) -- a tuning choice, same results */
} gl_plan_desc;

typedef struct gl_plan gl_plan;

/* Read-only facts about a built plan (for tooling / FLOP accounting). */
typedef struct gl_plan_info {
    int num_pix, supersample, n_params;
    int kernel_size;      /* sub-grid PSF size k */
    int folded_size;      /* kf = k + ss - 1 */
    int sep_rank;         /* > 0: separable PSF path with this rank, 0: direct 2-D path (fused kernel or,
                             for staged plans, their channel convolution) */
    int lds_bytes;        /* LDS per workgroup of the particle kernel */
    int block;            /* threads per workgroup */
    int adj_rank;         /* staged plans: rank of the separable adjoint conv, 0: direct (fused plans:
                             sep_rank covers both directions) */
} gl_plan_info;

/* Message of the last failing call on this thread. */
const char* gl_last_error(void);

/* Build an immutable device plan: copies PSF (pre-folded with the ss x ss
 * pooling box), observed image, prior table and component table to device
 * memory.  Replaces LensSimulator.__init__ (tf/simulator.py:11-55) plus the
 * constant state of ForwardProbModel/BackwardProbModel.__init__
 * (tf/model.py:29-51, 109-132). */
int gl_plan_create(const gl_plan_desc* desc, gl_plan** out);
int gl_plan_destroy(gl_plan* plan);
int gl_plan_query(const gl_plan* plan, gl_plan_info* info);

/* Batched forward model on physical parameters x (bs x d):
 * img (bs x N x N).  Replaces LensSimulator.simulate (tf/simulator.py:65-97). */
int gl_simulate(const gl_plan* plan, const float* x, int bs, float* img, hipStream_t stream);

/* Vector-Jacobian product of gl_simulate: grad_x (bs x d) = d<g_img, img>/dx.
 * Backward of simulate for autograd (the reference's GradientTape through
 * tf/simulator.py:65-97).  Recomputes the ray-trace; stores no grids. */
int gl_simulate_vjp(const gl_plan* plan, const float* x, int bs, const float* g_img,
                    float* grad_x, hipStream_t stream);

/* Log posterior on unconstrained z (bs x d): lp (bs), chi2 = mean r^2 (bs).
 * Replaces ForwardProbModel.log_prob (tf/model.py:53-86) and, for plans with
 * GL_LIK_BACKWARD and least-squares components, BackwardProbModel.log_prob
 * (tf/model.py:134-163).  If grad_z is non-NULL it receives
 * g_lp[i] * dlp_i/dz_i + g_chi2[i] * dchi2_i/dz_i (g_lp NULL: 1, g_chi2 NULL: 0),
 * computed by the hand-written adjoint in the same launch.  lp/chi2 may be NULL. */
int gl_log_prob(const gl_plan* plan, const float* z, int bs, float* lp, float* chi2,
                const float* g_lp, const float* g_chi2, float* grad_z, hipStream_t stream);

/* Least-squares simulation on physical parameters x (bs x d) of a plan whose light
 * components are linear (use_lstsq): every component is simulated with unit
 * amplitude, convolved and pooled, the amplitudes are solved per particle by
 * weighted least squares against the plan's observed image and noise map
 * (pseudo-inverse cutoff 1e-6 of the largest singular value, FP64 solve), and
 * the outputs are written when non-NULL: img (bs x N x N), coeffs (bs x D),
 * stacked (bs x N x N x D, the unit-amplitude components).  Replaces
 * LensSimulator.lstsq_simulate (tf/simulator.py:99-165) with its
 * return_stacked / return_coeffs outputs. */
int gl_lstsq_simulate(const gl_plan* plan, const float* x, int bs, float* img, float* coeffs,
                      float* stacked, hipStream_t stream);

/* gl_log_prob over several lens systems of one shape in one launch: particle i
 * is compared with observed image lens[i] of the plan (lens NULL: image 0).
 * An index outside [0, n_observed) gives NaN lp and chi2 for that particle.
 * No counterpart in the reference (which builds one simulator per lens). */
int gl_log_prob_lenses(const gl_plan* plan, const float* z, const int* lens, int bs, float* lp, float* chi2,
                       const float* g_lp, const float* g_chi2, float* grad_z, hipStream_t stream);

/* One multi-start MAP step with Keras-Adam semantics, fused with log_prob and
 * its adjoint: loss = mean_i(-lp_i / N^2) (tf/inference.py:28-35), then
 * m, v, z updated in place (tf.keras Adam, epsilon outside the sqrt).
 * step is 1-based.  lp / chi2 (bs) receive the pre-update values. */
int gl_map_adam_step(const gl_plan* plan, float* z, float* m, float* v, int bs, int step,
                     float lr, float beta1, float beta2, float epsilon,
                     float* lp, float* chi2, hipStream_t stream);

/* gl_map_adam_step over several lens systems of one shape in one launch
 * (lens as in gl_log_prob_lenses).  Particle i's loss term is
 * -weight[i] * lp_i / N^2 (weight NULL: 1/bs), so weight[i] = 1 / (particles of
 * its lens) reproduces independent per-lens MAP runs (tf/inference.py:28-35). */
int gl_map_adam_step_lenses(const gl_plan* plan, float* z, float* m, float* v, const int* lens,
                            const float* weight, int bs, int step, float lr, float beta1, float beta2,
                            float epsilon, float* lp, float* chi2, hipStream_t stream);

/* gl_map_adam_step without any device memory the library allocates: no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Resident-image plans only
 * (gl_plan_info.lds_bytes > 0; GL_ERR_UNSUPPORTED otherwise) and bs <= 16384
 * (GL_ERR_ARG beyond).  The launch-order buffer and the batch-coupled EPL series
 * counts come from a static device array of 16,392 ints (64 KiB) per GPU, held until
 * the process ends, instead of the plan's own buffers; results equal
 * gl_map_adam_step's.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (several streams or host threads) race and must be ordered by the caller. */
int gl_map_adam_step_noalloc(const gl_plan* plan, float* z, float* m, float* v, int bs, int step,
                             float lr, float beta1, float beta2, float epsilon,
                             float* lp, float* chi2, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Batches of lens systems with different image sizes, supersampling, PSFs and
 * particle counts.  No counterpart in the reference (one simulator per lens,
 * tf/inference.py:21).  Each lens is a fixed-amplitude plan with Sersic lights (same
 * parameter count and likelihood); the plans must outlive the batch.  Particles are
 * lens-major: lens 0's counts[0] particles first.  Lenses whose plan runs on the
 * resident-image kernel (gl_plan_info.lds_bytes > 0, at most 16384 particles) take
 * one fused launch each; the others run the staged pipeline, grouped into chunks
 * whose workspaces share one arena of at most workspace_bytes (0 selects 32 GiB; a
 * lens larger than that forms a chunk alone), every stage one launch over a chunk.
 * gl_batch_create is gl_batch_create_ws with workspace_bytes 0.
 * ------------------------------------------------------------------------- */
typedef struct gl_batch gl_batch;
int gl_batch_create(const gl_plan* const* plans, int n_lenses, const int* counts, gl_batch** out);
int gl_batch_create_ws(const gl_plan* const* plans, int n_lenses, const int* counts, size_t workspace_bytes,
                       gl_batch** out);
/* Options of gl_batch_create_ex.  band selects the banded persistent kernel: every
 * lens on it runs in ONE launch per call -- workgroups as many as the CUs hold at
 * once take the (lens, particle) items round-robin, longest lenses first, and
 * stream each particle's supersampled image through LDS in bands of sub-rows, so
 * images of any size share the launch.  It takes lenses of the profile combinations
 * it is instantiated for (EPL + shear with a Sersic ellipse source, with or without a
 * Sersic ellipse lens light) whose folded PSF has rank <= 8 at singular value 1e-6;
 * other lenses keep the resident-image or staged path.  band 0: off (gl_batch_create_ws);
 * 1: every lens it takes; 2: only lenses the resident-image kernel cannot hold.
 * band_rows caps the sub-rows per band (0: as many as the LDS holds, at most 64);
 * band_lds_bytes is the LDS per workgroup (0: 80 KiB less 256 bytes, two workgroups
 * per CU).  The kernel keeps R x nsub x N + N x N floats per workgroup in HBM
 * (allocated here). */
typedef struct gl_batch_opts {
    size_t workspace_bytes;  /* arena of the staged lenses (0: 32 GiB) */
    int band;
    int band_rows;
    int band_lds_bytes;
} gl_batch_opts;
int gl_batch_create_ex(const gl_plan* const* plans, int n_lenses, const int* counts, const gl_batch_opts* opts,
                       gl_batch** out);
typedef struct gl_batch_info {
    int total;          /* particles */
    int lenses_fused;   /* lenses on the resident-image kernel (one launch each) */
    int lenses_staged;  /* lenses on the staged pipeline */
    int lenses_band;    /* lenses on the banded persistent kernel (one launch for all) */
    int chunks;         /* staged chunks */
    int band_grid;      /* workgroups of the banded launch */
    int band_lds_bytes;
    int band_rows_min, band_rows_max;  /* sub-rows per band over its lenses */
    size_t band_scratch_bytes;         /* HBM of the banded kernel's per-workgroup slots */
    size_t workspace_bytes;            /* staged arena */
} gl_batch_info;
int gl_batch_query(const gl_batch* batch, gl_batch_info* info);
int gl_batch_destroy(gl_batch* batch);
int gl_batch_size(const gl_batch* batch);  /* total particles */
size_t gl_batch_workspace_bytes(const gl_batch* batch);  /* bytes of the staged lenses' arena */
int gl_batch_num_chunks(const gl_batch* batch);          /* staged chunks (launch sequences) per call */
/* lp, chi2 (total) and, when grad_z is non-NULL, dlp/dz (total x d) */
int gl_batch_log_prob(const gl_batch* batch, const float* z, float* lp, float* chi2, float* grad_z,
                      hipStream_t stream);
/* One MAP step of every lens at once: lens L's loss is mean over its particles of
 * -lp / N_L^2 (tf/inference.py:28-35) with Keras Adam; lp / chi2 get the
 * pre-update values. */
int gl_batch_map_adam_step(const gl_batch* batch, float* z, float* m, float* v, int step, float lr,
                           float beta1, float beta2, float epsilon, float* lp, float* chi2,
                           hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Diagonal-Gaussian SVI of many lens systems on the device (the surrogate of
 * tf/inference.py:43-73 restricted to a diagonal scale, as the uneven-batch
 * workload asks).  Lens L's surrogate is N(loc[L], diag(sigma^2)), sigma =
 * exp(raw[L]) + 1e-6 (FillScaleTriL's Exp bijector and diag_shift, :56-63), with
 * n_vi samples per lens laid out lens-major (particle L n_vi + s), i.e. a batch
 * whose counts are all n_vi.  One SVI step is gl_svi_diag_sample, the batch's
 * log_prob with gradient on z, then gl_svi_diag_update; no host work, so the
 * three calls can be recorded into a HIP graph and replayed.  step: per-lens
 * device counters (n_lenses ints, start at 0) that the update advances.
 * ------------------------------------------------------------------------- */
/* z, eps (n_lenses n_vi x d): eps ~ N(0, 1) from Philox4x32-10 keyed by seed with
 * counter (particle, dimension / 4, step[L], 0) and Box-Muller; z = loc + sigma eps. */
int gl_svi_diag_sample(const float* loc, const float* raw, int n_lenses, int n_vi, int d,
                       unsigned long long seed, const int* step, float* z, float* eps, hipStream_t stream);
/* Per lens, from lp (n_lenses n_vi) and grad_z = dlp/dz: the reparameterised
 * gradient of the loss mean_s [log q(z_s) - lp_s] (tfp.vi.fit_surrogate_posterior,
 * tf/inference.py:65-71) w.r.t. loc and raw, then Keras Adam (m, v: 2 n_lenses d,
 * [loc | raw]) with lr_t from step[L] + 1; step[L] += 1.  loss_hist (hist_len x
 * n_lenses, may be NULL) receives the loss at row step[L] (before the increment). */
int gl_svi_diag_update(const float* lp, const float* grad_z, const float* eps, int n_lenses, int n_vi, int d,
                       float* loc, float* raw, float* m, float* v, int* step, float lr, float beta1,
                       float beta2, float epsilon, float* loss_hist, int hist_len, hipStream_t stream);

/* Log posterior (and gradient) like gl_log_prob, as step k of a loop whose trip count
 * *limit lives in device memory: the launch does nothing when k >= *limit, so a
 * loop of data-dependent length can be recorded into a HIP graph at its maximum
 * length.  Resident-image plans only (GL_ERR_UNSUPPORTED otherwise).  No counterpart
 * in the reference (its HMC runs a tf.while_loop, tfp.mcmc). */
int gl_log_prob_masked(const gl_plan* plan, const float* z, int bs, float* lp, float* chi2,
                       const float* g_lp, const float* g_chi2, float* grad_z, const int* limit, int k,
                       hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Preconditioned HMC with dual-averaging step size and ChEES trajectory-length
 * adaptation on the device (the kernel stack of tf/inference.py:98-131).  n chains of
 * d <= 128 parameters.  State: si (4 ints: iteration, leapfrog steps L of the current
 * trajectory), sf (8 floats: step size, trajectory jitter, log T, dual-averaging
 * h_bar, log eps_bar, mu, averaged squared trajectory gradient, averaged log T).
 * One iteration:
 * gl_hmc_begin; for k < max_leapfrog: gl_hmc_position(k), target log-density and
 * gradient at z (gl_log_prob_masked with limit si + 1), gl_hmc_momentum(k);
 * gl_hmc_end; gl_hmc_adapt.  Steps k >= L change nothing, so the sequence can be
 * recorded once into a HIP graph.  Mass matrix: cov (d x d) with chol_inv its
 * inverse Cholesky factor; momenta p ~ N(0, cov^-1).
 * ------------------------------------------------------------------------- */
int gl_hmc_begin(int n, int d, const float* chol_inv, unsigned long long seed, const int* si, const float* sf,
                 const float* z, const float* lp, const float* g, float* p, float* z0, float* lp0, float* g0,
                 float* p0, hipStream_t stream);
int gl_hmc_position(int n, int d, const float* cov, const int* si, const float* sf, int k, float* z,
                    const float* p, hipStream_t stream);
int gl_hmc_momentum(int n, int d, const int* si, const float* sf, int k, float* p, const float* g,
                    hipStream_t stream);
/* Metropolis correction (energy -lp + p^T cov p / 2); proposal zp and velocity cov p
 * kept; acc (n) the acceptance probabilities; samples (n_results x n x d) from
 * iteration burnin on. */
int gl_hmc_end(int n, int d, const float* cov, unsigned long long seed, const int* si, float* z, float* lp,
               float* g, const float* p, const float* z0, const float* lp0, const float* g0, const float* p0,
               float* zp, float* vel, float* acc, float* samples, int burnin, int n_results,
               hipStream_t stream);
/* adapt = 0: fixed step size and L.  Histories (hist_len entries, by iteration): step
 * size, mean acceptance, L, trajectory length. */
int gl_hmc_adapt(int n, int d, int* si, float* sf, const float* z0, const float* zp, const float* vel,
                 const float* acc, int adapt, int n_adapt, float target, int max_leapfrog, float* eps_hist,
                 float* acc_hist, int* l_hist, float* t_hist, int hist_len, hipStream_t stream);

/* One leapfrog step of preconditioned HMC on a plan's log posterior (the integrator
 * of tfp PreconditionedHamiltonianMonteCarlo as used by tf/inference.py:109-117):
 * p += eps/2 grad; z += eps cov p; (lp, chi2, grad) at the new z; p += eps/2 grad.
 * z, p, grad_z (n x d, d = the plan's parameters): grad_z holds d lp / dz at z on
 * entry and at the new z on return.  cov (d x d) is the mass matrix's inverse
 * (kinetic energy p^T cov p / 2); NULL: identity. */
int gl_leapfrog_step(const gl_plan* plan, int n, const float* cov, float eps, float* z, float* p, float* lp,
                     float* chi2, float* grad_z, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Observation noise on the device (training images of the learned initialiser):
 * out[i] = Poisson(max(img[i], 0) exp_time) / exp_time + background_rms N(0, 1), the
 * noise model behind the likelihood variance of tf/model.py:73-76.  Draws from
 * Philox4x32-10 keyed by seed with counter (i, block, 0x5A17): a pure function of
 * (seed, i).  Poisson by CDF inversion below mean 10 and PTRS (Hoermann 1993) above.
 * img and out may be the same buffer.  No allocation; graph-capturable.
 * ------------------------------------------------------------------------- */
int gl_noise_poisson_gauss(const float* img, long long n, float exp_time, float background_rms,
                           unsigned long long seed, float* out, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Learned initialiser.  No counterpart in the reference: a small BF16 CNN
 * (hand-written MFMA kernels, FP32 accumulation) mapping an observed image to
 * a diagonal Gaussian (mean, log-std) over the unconstrained parameters, used
 * to seed ModellingSequence.MAP (tf/inference.py:17-41) next to prior draws.
 * Architecture: 4 x (3x3 conv pad 1 + bias + tanh-GELU), strides 1,2,1,2,
 * global average pool, Linear(c4 -> hidden) + GELU, Linear(hidden -> 2 n_out).
 * Parameters are one FP32 vector, per conv layer W[co][(kh*3+kw)*cin + ci]
 * then b[co], then W1[hidden][c4], b1, W2[2 n_out][hidden], b2.
 * ------------------------------------------------------------------------- */
typedef struct gl_cnn gl_cnn;

typedef struct gl_cnn_desc {
    int height, width;    /* input image */
    int channels[4];      /* conv widths, multiples of 32 */
    int hidden;           /* MLP width, <= 512 */
    int n_out;            /* d: parameters predicted (outputs 2 d), 2 d <= 256 */
    int max_batch;        /* images per call (workspace is sized for it) */
    float in_scale;       /* input transform asinh(img / in_scale) */
} gl_cnn_desc;

int gl_cnn_create(const gl_cnn_desc* desc, gl_cnn** out);
int gl_cnn_destroy(gl_cnn* net);
int gl_cnn_num_params(const gl_cnn* net);
/* Copy a parameter vector in (device memory) and reset the AdamW moments. */
int gl_cnn_set_params(gl_cnn* net, const float* params, hipStream_t stream);
int gl_cnn_get_params(const gl_cnn* net, float* params, hipStream_t stream);
/* out (bs x 2 n_out): mean, then log-std (unclamped) of each image (bs x H x W). */
int gl_cnn_forward(gl_cnn* net, const float* images, int bs, float* out, hipStream_t stream);
/* Mean Gaussian NLL over the batch against target (bs x n_out), log-std clamped to
 * [-7, 5]; its gradient is kept for gl_cnn_adamw_step and copied to grad when
 * non-NULL.  loss: one float (device), may be NULL. */
int gl_cnn_loss_grad(gl_cnn* net, const float* images, const float* target, int bs, float* loss, float* grad,
                     hipStream_t stream);
/* AdamW (torch.optim.AdamW semantics, step 1-based) on the kept gradient. */
int gl_cnn_adamw_step(gl_cnn* net, int step, float lr, float beta1, float beta2, float eps,
                      float weight_decay, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif

/*
 * llgmi.h -- C ABI of the MI355X (gfx950) LLG hot-path library (libllgmi.so).
 *
 * Drop-in boundary for llg-sim's uniform-grid path.  Each entry point names
 * the Rust function it replaces (file:line in the llg-sim source tree).  The
 * host-side entry points take the reference's own layout: AoS f64 vector
 * fields, one [x, y, z] triple per cell, cell index j*nx + i
 * (grid.rs:54-58, vector_field.rs:24-27), so a Rust `extern "C"` block can
 * pass `VectorField2D::data.as_ptr()` straight through.
 *
 * Conventions
 *  - Every call returns an int status: LLGMI_OK (0) or an error code; the
 *    message of the last failure on the calling thread is llgmi_last_error().
 *    Nothing aborts the process; invalid arguments are rejected before any
 *    device work is queued.
 *  - `add_*` functions accumulate into the output field, `compute_*` and
 *    `build_*` overwrite it, exactly as the Rust names do.
 *  - A context binds one grid, one precision and one GPU.  It owns the
 *    cached demag operators (the reference's process-global DEMAG_CACHE,
 *    demag_fft_uniform.rs:44-108, and DEMAG_MG_CACHE, demag_poisson_mg.rs:3603),
 *    the RK45 scratch and FSAL state (RK45Scratch, llg.rs:126-189), and one HIP
 *    stream.  A context is not thread-safe: callers serialise on it, as the
 *    reference serialises on each operator's Mutex.
 *  - Host entry points are synchronous.  Device entry points (llgmi_dev_*)
 *    take SoA device buffers in the context's precision (component-major:
 *    x[0..n), y[0..n), z[0..n)), are ordered on the context stream and return
 *    once the work is queued, except where a host-visible result is returned
 *    (the RK45 step returns eps/accepted/dt_used and therefore synchronises).
 */
#ifndef LLGMI_H
#define LLGMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define LLGMI_OK 0
#define LLGMI_ERR_INVALID 1     /* bad argument (null pointer, size, enum) */
#define LLGMI_ERR_HIP 2         /* HIP runtime error (message has the HIP text) */
#define LLGMI_ERR_UNSUPPORTED 3 /* valid for the reference, not built here */
#define LLGMI_ERR_ALLOC 4       /* device allocation failed */

/* ---- enums ---------------------------------------------------------------- */
/* FieldMask (effective_field/mod.rs:37-45) */
#define LLGMI_MASK_EXCH_ANIS 0
#define LLGMI_MASK_EXCH_ANIS_DMI 1
#define LLGMI_MASK_FULL 2
/* DemagMethod (params.rs:67-92) */
#define LLGMI_DEMAG_FFT 0
#define LLGMI_DEMAG_MG 1
#define LLGMI_DEMAG_DST 2 /* retired in the reference: routed to MG (demag.rs:128-141) */
/* arithmetic of a context */
#define LLGMI_F32 0
#define LLGMI_F64 1

/* ---- parameter structs (plain data, 8-byte aligned) ------------------------ */
/* Grid2D (grid.rs:5-31) */
typedef struct llgmi_grid {
    int64_t nx, ny;
    double dx, dy, dz;
} llgmi_grid;

/* Material (params.rs:94-109).  `dmi` is Option<f64>: has_dmi = 0 is None. */
typedef struct llgmi_material {
    double ms;            /* A/m */
    double a_ex;          /* J/m */
    double k_u;           /* J/m^3 */
    double easy_axis[3];
    int32_t has_dmi;
    int32_t demag;        /* include the demag field */
    double dmi;           /* J/m^2 */
    int32_t demag_method; /* LLGMI_DEMAG_* */
    int32_t reserved;
} llgmi_material;

/* LLGParams (params.rs:137-144) */
typedef struct llgmi_params {
    double gamma;    /* rad/(s T) */
    double alpha;
    double dt;       /* s; in/out for the RK45 step */
    double b_ext[3]; /* T */
} llgmi_params;

typedef struct llgmi_ctx llgmi_ctx;

/* ---- library ------------------------------------------------------------- */
const char *llgmi_last_error(void);
int llgmi_version(void);
/* Number of HIP devices visible (0 when no GPU). */
int llgmi_device_count(int *count);

/* ---- context ---------------------------------------------------------------- */
/* Creates a context for `grid` in `precision` (LLGMI_F32 / LLGMI_F64) on HIP
 * device `device`.  Replaces RK45Scratch::new (llg.rs:159) plus the lazily
 * built operator caches. */
int llgmi_ctx_create(const llgmi_grid *grid, int precision, int device, llgmi_ctx **out);
int llgmi_ctx_destroy(llgmi_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own
 * work or events against the library's kernels. */
int llgmi_ctx_stream(llgmi_ctx *ctx, void **hip_stream);
int llgmi_ctx_synchronize(llgmi_ctx *ctx);

/* ---- host entry points (AoS f64, synchronous) -------------------------------- */
/* build_h_eff_masked (effective_field/mod.rs:48-75): overwrite b_eff with
 * Zeeman + exchange + anisotropy (+ DMI) (+ demag). */
int llgmi_build_h_eff_masked(llgmi_ctx *ctx, const double *m, double *b_eff,
                             const llgmi_params *params, const llgmi_material *mat,
                             int field_mask);
/* build_h_eff_masked_geom (effective_field/mod.rs:78-101); geom_mask may be
 * NULL (= None); one byte per cell, nonzero = material. */
int llgmi_build_h_eff_masked_geom(llgmi_ctx *ctx, const double *m, double *b_eff,
                                  const llgmi_params *params, const llgmi_material *mat,
                                  int field_mask, const uint8_t *geom_mask);
/* build_h_eff (effective_field/mod.rs:104-112): FieldMask::Full. */
int llgmi_build_h_eff(llgmi_ctx *ctx, const double *m, double *b_eff,
                      const llgmi_params *params, const llgmi_material *mat);

/* demag::add_demag_field_pbc (effective_field/demag.rs:70-143): accumulate.
 * Dispatch follows mat->demag_method (MG for open boundaries, FFT when
 * pbc_x or pbc_y > 0 or when the method is FFT).  pbc > 0 makes that axis
 * periodic with the MuMax-style image sums of build_kernel_realspace_pbc
 * (demag_fft_uniform.rs:941-1005); the periodic operator is cached in the
 * context beside the open one and rebuilt when pbc_x / pbc_y change.  Image
 * ranges beyond 2^27 kernel displacements return LLGMI_ERR_UNSUPPORTED. */
int llgmi_add_demag_field_pbc(llgmi_ctx *ctx, const double *m, double *b_eff,
                              const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y);
/* demag::compute_demag_field_pbc (effective_field/demag.rs:146-156): overwrite. */
int llgmi_compute_demag_field_pbc(llgmi_ctx *ctx, const double *m, double *out,
                                  const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y);
/* demag::add_demag_field (effective_field/demag.rs:47-55), open boundaries. */
int llgmi_add_demag_field(llgmi_ctx *ctx, const double *m, double *b_eff,
                          const llgmi_material *mat);

/* step_llg_rk45_recompute_field_masked_adaptive (llg.rs:795-1026).
 * m (AoS f64) is advanced in place when the step is accepted; params->dt is
 * replaced by the next step size; eps, accepted (0/1) and dt_used are the
 * Rust tuple (f64, bool, f64).  FSAL state lives in the context. */
int llgmi_rk45_step(llgmi_ctx *ctx, double *m, llgmi_params *params,
                    const llgmi_material *mat, int field_mask, double max_err,
                    double headroom, double dt_min, double dt_max, double *eps,
                    int *accepted, double *dt_used);
/* step_llg_rk45_recompute_field_adaptive (llg.rs:1029-1050): FieldMask::Full. */
int llgmi_rk45_step_adaptive(llgmi_ctx *ctx, double *m, llgmi_params *params,
                             const llgmi_material *mat, double max_err, double headroom,
                             double dt_min, double dt_max, double *eps, int *accepted,
                             double *dt_used);
/* Drops the FSAL cache (what a fresh RK45Scratch::new gives). */
int llgmi_rk45_reset(llgmi_ctx *ctx);

/* ---- device-resident entry points (SoA, context precision) -------------------- */
/* Same as llgmi_build_h_eff_masked_geom, on device buffers of 3*n elements.
 * geom_mask_dev: device bytes or NULL.  Asynchronous on the context stream. */
int llgmi_dev_build_h_eff(llgmi_ctx *ctx, const void *m_dev, void *b_dev,
                          const llgmi_params *params, const llgmi_material *mat,
                          int field_mask, const uint8_t *geom_mask_dev);
/* llgmi_dev_build_h_eff that allocates nothing (no hipMalloc, no library
 * allocation), for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation.  No
 * counterpart in the reference.  Its scratch (the FFT spectrum, plus B_demag
 * when the inverse row pass runs unfused) is one 48 MiB device array per GPU,
 * held until the process ends: grids whose scratch exceeds it (FP32 up to
 * 1024 x 1024 fits, FP64 up to 1024 x 512) get LLGMI_ERR_UNSUPPORTED.  The
 * context's FFT operator must already be built and tuned by one
 * llgmi_dev_build_h_eff call with the Full mask (LLGMI_ERR_INVALID
 * otherwise); the multigrid demag is not offered (LLGMI_ERR_UNSUPPORTED).
 * All calls on one GPU share the array: calls that overlap on a GPU (other
 * streams, other threads) race. */
int llgmi_dev_build_h_eff_noalloc(llgmi_ctx *ctx, const void *m_dev, void *b_dev,
                                  const llgmi_params *params, const llgmi_material *mat,
                                  int field_mask, const uint8_t *geom_mask_dev);
/* Same as llgmi_add_demag_field_pbc on device buffers (accumulate). */
int llgmi_dev_add_demag_field_pbc(llgmi_ctx *ctx, const void *m_dev, void *b_dev,
                                  const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y);
/* Same as llgmi_rk45_step with the state in device memory (SoA). */
int llgmi_dev_rk45_step(llgmi_ctx *ctx, void *m_dev, llgmi_params *params,
                        const llgmi_material *mat, int field_mask, double max_err,
                        double headroom, double dt_min, double dt_max, double *eps,
                        int *accepted, double *dt_used);
/* llgmi_dev_rk45_step for `count` independent films at once: film i is
 * ctxs[i] (distinct contexts on one device, same precision) with state
 * m_devs[i], params[i] (dt in/out, per-film adaptive) and mats[i]; eps,
 * accepted and dt_used are arrays of `count`.  With FFT demag (or no demag)
 * on every film the attempt runs as batched kernels, one launch per kernel
 * class over all films replayed from one HIP graph, with one host
 * synchronisation for all films' controllers; otherwise every film's attempt
 * is launched on its own stream before any is waited for.  Each film's
 * result is what llgmi_dev_rk45_step gives for it alone.  The reference has
 * no batched stepper: its drivers call llg.rs:795-1026 once per film, which
 * is what this replaces for a loop over films.  On an error no film's state
 * or dt changes. */
int llgmi_dev_rk45_step_batch(llgmi_ctx *const *ctxs, void *const *m_devs, llgmi_params *params,
                              const llgmi_material *mats, int64_t count, int field_mask,
                              double max_err, double headroom, double dt_min, double dt_max,
                              double *eps, int *accepted, double *dt_used);
/* `attempts` consecutive llgmi_dev_rk45_step attempts of each of `count`
 * films with the controller (llg.rs:999-1025, FP64) on the device: every
 * attempt's eps, decision and step size are formed there, the next attempt
 * starts from them (FSAL on accept), and the host waits once for the whole
 * run.  eps, accepted and dt_used are attempt-major arrays of
 * attempts * count ([a * count + i]); params[i].dt returns the step size for
 * the film's next attempt.  attempts <= 256.  Films as for
 * llgmi_dev_rk45_step_batch, with the FFT demag or none
 * (LLGMI_ERR_UNSUPPORTED otherwise, and for FP64 films with ny > 2048).
 * Results are those of the same attempts taken one call at a time up to the
 * last-ulp rounding of the controller's power function (device against host
 * libm): a step size an ulp apart moves the following error norms by up to
 * ~1e-10 relative (the norm cancels to ~1e-4 of the slopes).  No reference
 * counterpart: it
 * replaces a driver's loop over llg.rs:795-1026 for callers that do not need
 * to look at each attempt before the next. */
int llgmi_dev_rk45_run_batch(llgmi_ctx *const *ctxs, void *const *m_devs, llgmi_params *params,
                             const llgmi_material *mats, int64_t count, int64_t attempts, int field_mask,
                             double max_err, double headroom, double dt_min, double dt_max, double *eps,
                             int *accepted, double *dt_used);
/* llgmi_dev_rk45_run_batch for one film. */
int llgmi_dev_rk45_run(llgmi_ctx *ctx, void *m_dev, llgmi_params *params, const llgmi_material *mat,
                       int64_t attempts, int field_mask, double max_err, double headroom, double dt_min,
                       double dt_max, double *eps, int *accepted, double *dt_used);
/* m = 0 on the cells whose keep byte (device, n bytes) is 0 (synchronous):
 * resets the holes of a geometry after a step whose stage normalisation
 * (normalize(0) = +z, vec3.rs:21-28) filled them. */
int llgmi_dev_mask_cells(llgmi_ctx *ctx, void *m_dev, const uint8_t *keep_dev);
/* AoS f64 host <-> SoA device copies in the context precision (synchronous). */
int llgmi_dev_upload(llgmi_ctx *ctx, const double *host_aos, void *dev_soa);
int llgmi_dev_download(llgmi_ctx *ctx, const void *dev_soa, double *host_aos);

/* ---- multigrid Poisson demag (demag_poisson_mg.rs) ------------------------------ */
/* DemagPoissonMGConfig (demag_poisson_mg.rs:115-185) together with the
 * operator choice MGOperatorSettings (:727-804).  144 bytes, no padding. */
typedef struct llgmi_mg_config {
    int64_t pad_xy;          /* vacuum cells on each side in x and y */
    int64_t n_vac_z;         /* vacuum layers below and above the film */
    int64_t v_cycles;        /* minimum V-cycles (exact count without a tolerance) */
    int64_t v_cycles_max;
    int32_t has_tol_abs;     /* Option<f64> tol_abs: 0 = None */
    int32_t has_tol_rel;     /* Option<f64> tol_rel: 0 = None */
    double tol_abs;          /* max-norm residual bound (A/m^2) */
    double tol_rel;          /* relative to max |rhs| */
    int64_t pre_smooth, post_smooth;
    int32_t smoother;        /* 0 weighted Jacobi, 1 red-black SOR (MGSmoother :97-112) */
    int32_t warm_start;      /* 0/1 */
    double omega;            /* weighted-Jacobi factor */
    double sor_omega;        /* red-black SOR factor */
    int32_t bc;              /* 0 Dirichlet zero, 1 monopole+dipole, 2 treecode (:81-94) */
    int32_t stencil;         /* 0 seven-point, 1 iso9+z, 2 iso27 (:708-712) */
    double tree_theta;
    int64_t tree_leaf;
    int64_t tree_max_depth;
    int32_t prolong;         /* 0 injection, 1 trilinear (:714-718) */
    int32_t coarse_op;       /* 0 rediscretize, 1 Galerkin (:720-724) */
    double iso27_alpha;      /* > 0: LLG_DEMAG_MG_ISO27_FLUX_ALPHA override; else none */
} llgmi_mg_config;

/* Outcome of the last MG solve of a context. */
typedef struct llgmi_mg_stats {
    int64_t cycles;          /* V-cycles run */
    double rhs_max;          /* max |rhs| over interior cells (0 without tol_rel) */
    double tol_target;       /* max(tol_abs, tol_rel * rhs_max) */
    double last_residual;    /* max |r| at the last check, -1 if none was made */
    int32_t diverged;        /* the NaN/Inf guard reset phi (:2021-2032) */
    int32_t n_charges;       /* treecode point charges */
    int32_t tree_rebuilt;    /* the treecode octree was rebuilt for this solve */
    int32_t levels;          /* grid levels of the hierarchy */
    int64_t box[3];          /* finest padded box px, py, pz */
} llgmi_mg_stats;

/* HybridConfig (demag_poisson_mg.rs:274-361) with the boundary-scaling knobs
 * LLG_DEMAG_MG_HYBRID_BSCALE / _BMARGIN (:2743-2760): the MG field plus a local
 * correction stencil dK = K_FFT - K_MG of radius radius_xy built from impulse
 * responses (:3159-3597), with the rhs screened by a Gaussian of sigma_cells.
 * Active when enabled != 0 and radius_xy > 0.  48 bytes. */
typedef struct llgmi_mg_hybrid_config {
    int32_t enabled;
    int32_t reserved;
    int64_t radius_xy;        /* default 12 */
    int64_t delta_v_cycles;   /* V-cycles of each impulse solve, default 60 */
    double sigma_cells;       /* default 1.5; 0 = unscreened */
    double bscale;            /* > 0: boundary scaling exponent; default 0 */
    int64_t bmargin;          /* default 2 */
} llgmi_mg_hybrid_config;

/* HybridConfig::default() (disabled). */
int llgmi_mg_hybrid_config_default(llgmi_mg_hybrid_config *cfg);
/* HybridConfig::from_env (:296-356) + the BSCALE / BMARGIN variables. */
int llgmi_mg_hybrid_config_from_env(llgmi_mg_hybrid_config *cfg);
/* Pins the hybrid configuration of a context; NULL reads the environment at
 * every MG call, as the reference does.  The correction stencil is cached in
 * the context and rebuilt when the MG or hybrid settings change. */
int llgmi_ctx_set_mg_hybrid_config(llgmi_ctx *ctx, const llgmi_mg_hybrid_config *cfg);

/* DemagPoissonMGConfig::default() with the default operator (iso27 stencil,
 * trilinear prolongation, Galerkin coarse operators). */
int llgmi_mg_config_default(llgmi_mg_config *cfg);
/* DemagPoissonMGConfig::from_env (:188-267) + MGOperatorSettings::from_env
 * (:741-794): the LLG_DEMAG_MG_* variables, with the reference's aliases,
 * clamps and fallbacks. */
int llgmi_mg_config_from_env(llgmi_mg_config *cfg);
/* Pins the MG configuration of a context.  NULL restores the reference
 * behaviour: the configuration is read from the environment at every call
 * (add_demag_field_poisson_mg, :3612-3613).  A smoother that the operator
 * does not allow is overridden as sanitize_cfg_for_op does (:1406-1417). */
int llgmi_ctx_set_mg_config(llgmi_ctx *ctx, const llgmi_mg_config *cfg);
/* add_demag_field_poisson_mg (demag_poisson_mg.rs:3605-3635): accumulate the
 * MG demag field; the solver (hierarchy, warm-start potential, treecode) is
 * cached in the context and rebuilt when grid, padding or operator change. */
int llgmi_add_demag_field_poisson_mg(llgmi_ctx *ctx, const double *m, double *b_eff,
                                     const llgmi_material *mat);
/* compute_demag_field_poisson_mg (:3638-3646): overwrite. */
int llgmi_compute_demag_field_poisson_mg(llgmi_ctx *ctx, const double *m, double *out,
                                         const llgmi_material *mat);
int llgmi_mg_last_stats(llgmi_ctx *ctx, llgmi_mg_stats *out);
/* The finest multigrid box of the context's last solve: dims[0..5] =
 * (px, py, pz, offx, offy, offz), film cell (i, j) at box (offx+i, offy+j,
 * offz).  No reference counterpart: these three serve the learned initial
 * guess below. */
int llgmi_mg_box(llgmi_ctx *ctx, int64_t dims[6]);
/* Copies the finest potential phi[z][y][x] (px*py*pz doubles, device memory)
 * out of, or into, the solver.  With warm_start on, a solve starts from the
 * phi written here (the boundary faces are restamped by the solve). */
int llgmi_mg_phi_read(llgmi_ctx *ctx, double *phi_dev);
int llgmi_mg_phi_write(llgmi_ctx *ctx, const double *phi_dev);

/* ---- learned initial guess for the multigrid solve (no reference counterpart) ---
 * A 2-D U-Net (4 levels, base_channels * {1,2,4,8} channels, bf16 weights and
 * activations on MFMA with FP32 accumulation, FP32 master weights and Adam)
 * mapping per-pixel inputs x[b][y][x][16] (bf16; (m_x, m_y, m_z, mask) in
 * channels 0..3) to outputs y[b][y][x][c_pad] (FP32, c_pad = out_channels
 * rounded up to 16); trained on the device.  The loss is the mean squared
 * error over channels < out_channels and pixels x < valid_width,
 * y < valid_height, against bf16 targets t[b][y][x][c_pad]. */
typedef struct llgmi_ws llgmi_ws;
typedef struct llgmi_ws_shape {
    int64_t batch, height, width;       /* height, width multiples of 8 */
    int32_t in_channels, out_channels;  /* 1..16, 1..256 */
    int32_t base_channels, reserved;    /* 16 or 32 */
    int64_t valid_width, valid_height;
} llgmi_ws_shape;                       /* 56 bytes */
int llgmi_ws_create(const llgmi_ws_shape *shape, uint64_t seed, int device, llgmi_ws **out);
int llgmi_ws_destroy(llgmi_ws *ws);
int llgmi_ws_param_count(llgmi_ws *ws, int64_t *n);
/* FP32 parameters in layer order (enc 0..3 x 2, dec 2..0 x 2, head), each
 * layer weights [tap][c_out][c_in] (tap = (dy+1)*3 + dx+1) then bias [c_out].
 * set_params also resets the Adam state. */
int llgmi_ws_get_params(llgmi_ws *ws, float *host);
int llgmi_ws_set_params(llgmi_ws *ws, const float *host);
int llgmi_ws_get_grads(llgmi_ws *ws, float *host);   /* of the last loss_grad */
int llgmi_ws_forward(llgmi_ws *ws, const uint16_t *x_dev, float *y_dev);
int llgmi_ws_loss_grad(llgmi_ws *ws, const uint16_t *x_dev, const uint16_t *t_dev, double *loss);
int llgmi_ws_adam_step(llgmi_ws *ws, double lr);
int llgmi_ws_train_step(llgmi_ws *ws, const uint16_t *x_dev, const uint16_t *t_dev, double lr,
                        double *loss);
/* Packing between the multigrid context and the network's tensors, for
 * sample `sample` of a batch: the film's m (SoA, FP32 or FP64, nx x ny) at
 * offset (offx, offy) of the plane; phi[z][y][x] of a px x py x pz box as
 * target channels z (divided by scale) and back (multiplied by scale; with
 * accumulate != 0 added to phi instead of overwriting it). */
int llgmi_ws_pack_input(llgmi_ws *ws, int64_t sample, const void *m_dev, int m_is_f32,
                        int64_t nx, int64_t ny, int64_t offx, int64_t offy, uint16_t *x_dev);
int llgmi_ws_pack_phi(llgmi_ws *ws, int64_t sample, const double *phi_dev, int64_t px,
                      int64_t py, int64_t pz, double scale, uint16_t *t_dev);
int llgmi_ws_unpack_phi(llgmi_ws *ws, int64_t sample, const float *y_dev, int64_t px,
                        int64_t py, int64_t pz, double scale, int accumulate, double *phi_dev);

/* ---- relaxation (llg.rs:38-50, 196-436; energy.rs; relax.rs) -------------------- */
/* TorqueMetric (relax.rs:32-38) */
#define LLGMI_TORQUE_MAX 0
#define LLGMI_TORQUE_MEAN 1
#define LLGMI_TORQUE_RMS 2
#define LLGMI_TORQUE_MEANSQ 3
/* RelaxStopReason (relax.rs:40-46) */
#define LLGMI_STOP_MAX_ACCEPTED_STEPS 0
#define LLGMI_STOP_PHASE2_DISABLED 1
#define LLGMI_STOP_TIGHTEN_FLOOR_REACHED 2
#define LLGMI_STOP_TORQUE_GATE_NOT_SATISFIED 3
/* RelaxStageStop (relax.rs:48-53); NONE = None */
#define LLGMI_STAGE_NONE (-1)
#define LLGMI_STAGE_BELOW_THRESHOLD 0
#define LLGMI_STAGE_PLATEAU 1
#define LLGMI_STAGE_MAX_ACCEPTED_STEPS 2

/* RelaxSettings (relax.rs:73-105); has_* = 0 is None for the Option fields.
 * 176 bytes, no padding. */
typedef struct llgmi_relax_settings {
    double max_err, headroom, dt_min, dt_max;
    int32_t phase1_enabled, phase2_enabled;
    int64_t energy_stride;
    double rel_energy_tol;
    int32_t torque_metric;            /* LLGMI_TORQUE_* */
    int32_t has_torque_threshold;
    double torque_threshold;
    int64_t torque_check_stride;
    double torque_plateau_rel, torque_plateau_abs;
    int64_t torque_plateau_checks, torque_plateau_min_checks;
    double tighten_factor, tighten_floor;
    int64_t max_accepted_steps;
    int32_t has_final_torque_gate, has_final_torque_gate_max;
    double final_torque_gate, final_torque_gate_max;
    int64_t gate_max_extra_accepted_steps, gate_plateau_fails;
} llgmi_relax_settings;

/* RelaxReport (relax.rs:55-71).  80 bytes. */
typedef struct llgmi_relax_report {
    int64_t accepted_steps, rejected_steps, torque_checks, torque_field_rebuilds;
    int32_t stop_reason;              /* LLGMI_STOP_* */
    int32_t last_stage_stop;          /* LLGMI_STAGE_* */
    int32_t has_final_torque, has_final_torque_max;
    double final_torque, final_torque_max, final_max_err, final_dt;
} llgmi_relax_report;

/* EnergyBreakdown (energy.rs:21-34), joules. */
typedef struct llgmi_energy {
    double exchange, anisotropy, zeeman, dmi, demag;
} llgmi_energy;

/* RelaxSettings::default() (relax.rs:107-139). */
int llgmi_relax_settings_default(llgmi_relax_settings *s);
/* mumax_like_relax_settings (relax.rs:789-808). */
int llgmi_relax_settings_mumax_like(llgmi_relax_settings *s);

/* step_llg_rk23_recompute_field_masked_relax_adaptive (llg.rs:278-436):
 * damping-only Bogacki-Shampine RK23 step with FSAL.  m (AoS f64) advances in
 * place when accepted; params->dt becomes the next step size.  The RK23
 * scratch and its FSAL / last-field state (RK23Scratch, llg.rs:196-268) live
 * in the context, apart from the RK45 state. */
int llgmi_rk23_relax_step(llgmi_ctx *ctx, double *m, llgmi_params *params,
                          const llgmi_material *mat, int field_mask, double max_err,
                          double headroom, double dt_min, double dt_max, double *eps,
                          int *accepted, double *dt_used);
/* RK23Scratch::invalidate_last_b_eff (llg.rs:264-268). */
int llgmi_rk23_invalidate(llgmi_ctx *ctx);
/* RK23Scratch::last_b_eff (llg.rs:252-258): *valid = 1 and b_out (AoS f64)
 * filled when the field of the accepted state is held; *valid = 0 otherwise. */
int llgmi_rk23_last_b_eff(llgmi_ctx *ctx, double *b_out, int *valid);
/* compute_energy_geom (energy.rs:80-283); geom_mask NULL = None.  The demag
 * term evaluates the demag field through the same dispatch as
 * llgmi_add_demag_field. */
int llgmi_compute_energy_geom(llgmi_ctx *ctx, const double *m, const llgmi_material *mat,
                              const double b_ext[3], const uint8_t *geom_mask,
                              llgmi_energy *out);
/* relax_with_report (relax.rs:334-774): m (AoS f64) and params->dt in place;
 * settings->max_err is written back as the reference does. */
int llgmi_relax_with_report(llgmi_ctx *ctx, double *m, llgmi_params *params,
                            const llgmi_material *mat, int field_mask,
                            llgmi_relax_settings *settings, llgmi_relax_report *report);
/* Device-resident forms (SoA, context precision; geom_mask_dev on the device). */
int llgmi_dev_rk23_relax_step(llgmi_ctx *ctx, void *m_dev, llgmi_params *params,
                              const llgmi_material *mat, int field_mask, double max_err,
                              double headroom, double dt_min, double dt_max, double *eps,
                              int *accepted, double *dt_used);
int llgmi_dev_compute_energy_geom(llgmi_ctx *ctx, const void *m_dev, const llgmi_material *mat,
                                  const double b_ext[3], const uint8_t *geom_mask_dev,
                                  llgmi_energy *out);
int llgmi_dev_relax_with_report(llgmi_ctx *ctx, void *m_dev, llgmi_params *params,
                                const llgmi_material *mat, int field_mask,
                                llgmi_relax_settings *settings, llgmi_relax_report *report);

/* ---- on-disk formats (ovf.rs; demag_fft_uniform.rs:206-429) ----------------------- */
/* OvfMeta (ovf.rs:17-46).  NULL fields (and a NULL meta) take
 * OvfMeta::magnetization(): title "m", labels m_x m_y m_z, units 1. */
typedef struct llgmi_ovf_meta {
    const char *title;
    const char *const *desc_lines;   /* n_desc lines, each written as "# Desc: <line>" */
    int32_t n_desc;
    int32_t reserved;
    const char *valuelabels[3];
    const char *valueunits[3];
} llgmi_ovf_meta;

/* write_ovf2_rectangular_text (ovf.rs:54-159): OVF 2.0 text, the reference's
 * header lines and Rust number formatting; parent directories are created.
 * m is AoS f64 (nx*ny cells). */
int llgmi_write_ovf2_text(const char *path, const llgmi_grid *grid, const double *m,
                          const llgmi_ovf_meta *meta);
/* write_ovf2_rectangular_binary4 (ovf.rs:161-262): OVF 2.0 binary4 (f32 LE,
 * check value 1234567.0). */
int llgmi_write_ovf2_binary4(const char *path, const llgmi_grid *grid, const double *m,
                             const llgmi_ovf_meta *meta);
/* Either writer on a device-resident field (SoA, context precision). */
int llgmi_dev_write_ovf2(llgmi_ctx *ctx, const void *m_dev, const char *path,
                         const llgmi_ovf_meta *meta, int binary4);
/* OvfMeta::with_total_sim_time (ovf.rs:36-41): the desc line for time t_s,
 * NUL-terminated into buf (len bytes). */
int llgmi_ovf_total_sim_time_line(double t_s, char *buf, size_t len);

/* demag_cache_path (demag_fft_uniform.rs:262-280): the reference's cache file
 * name, relative to the working directory, into buf (len bytes). */
int llgmi_demag_cache_path(const llgmi_grid *grid, uint64_t pbc_x, uint64_t pbc_y, char *buf,
                           size_t len);
/* write_kernel_kspace (demag_fft_uniform.rs:282-320): writes the context's FFT
 * demag kernel (built first if needed) as an LLGDMAG v2 file at the reference's
 * FFT sizes; path NULL = llgmi_demag_cache_path. */
int llgmi_demag_cache_save(llgmi_ctx *ctx, uint64_t pbc_x, uint64_t pbc_y, const char *path);
/* try_load_kernel_kspace (demag_fft_uniform.rs:322-429): builds the context's
 * FFT demag operator from an LLGDMAG v2 file (no quadrature).  *loaded = 0 and
 * nothing changes when the file is missing or its header does not match
 * (grid, pbc, sizes, accuracy), as in the reference.  With LLGMI_DEMAG_CACHE=1
 * in the environment every operator build loads and saves the reference's
 * cache file itself, as Demag2D::new does. */
int llgmi_demag_cache_load(llgmi_ctx *ctx, uint64_t pbc_x, uint64_t pbc_y, const char *path,
                           int *loaded);

/* ---- per-kernel timing (HIP events on the context stream) ---------------------- */
/* Kernel classes timed when profiling is on. */
#define LLGMI_K_FFT_ROW_FWD 0   /* row R2C of Ms*m (zero padded) */
#define LLGMI_K_FFT_COL 1       /* column FFT, k-space tensor multiply, inverse column FFT */
#define LLGMI_K_FFT_ROW_INV 2   /* row C2R to B_demag */
#define LLGMI_K_LOCAL_FIELD 3   /* Zeeman + exchange + anisotropy + DMI (+ torque) */
#define LLGMI_K_RK45_STAGE 4    /* stage state combination + normalise */
#define LLGMI_K_RK45_ERR 5      /* error-norm reduction */
#define LLGMI_K_MG 6            /* multigrid RHS, smoothing, transfers, residuals, gradient */
#define LLGMI_K_MG_BC 7         /* multigrid boundary values (treecode / dipole) */
#define LLGMI_K_RELAX 8         /* RK23 stage states, error norm, energy and torque reductions */
#define LLGMI_K_COUNT 9
/* Times the kernel classes whose bit (1 << class) is set in class_mask; 0 stops. */
int llgmi_profile_enable(llgmi_ctx *ctx, int class_mask);
int llgmi_profile_reset(llgmi_ctx *ctx);
/* Sum of event-timed durations (ms) and number of launches of a class. */
int llgmi_profile_read(llgmi_ctx *ctx, int kernel_class, double *total_ms, int64_t *count);
/* Average duration (us) of each FFT demag pass of the context's open-boundary
 * operator, launched `reps` times back to back between two events (no event
 * between launches): avg_us[0] row R2C of ms*m_dev, avg_us[1] column pass (in
 * place on the spectrum, whose values it scales down each time), avg_us[2]
 * row C2R into the context's demag buffer.  Leaves the spectrum changed; the
 * next field evaluation writes it again.  Synchronous.  Measurement aid with
 * no counterpart in the reference (which times whole calls,
 * bin/demag_fft_vs_mg.rs:229-240). */
int llgmi_profile_fft_passes(llgmi_ctx *ctx, const void *m_dev, double ms, int reps, double *avg_us);

#ifdef __cplusplus
}
#endif
#endif

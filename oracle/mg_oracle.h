/*
 * mg_oracle.h -- CPU FP64 restatement of llg-sim's multigrid Poisson demag
 * (demag_poisson_mg.rs, mg_kernels.rs, mg_treecode.rs).
 *
 * TEST INFRASTRUCTURE ONLY, like llg_oracle.h: the checker for the HIP
 * multigrid path and the CPU baseline beside it.  The product never links it.
 *
 * Box arrays are FP64, index (k*ny + j)*nx + i (mg_kernels.rs:20-22).
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H

#include <stdint.h>

#include "llg_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DemagPoissonMGConfig (demag_poisson_mg.rs:115-185) plus MGOperatorSettings
 * (:727-804).  Same field order and sizes as llgmi_mg_config (144 bytes). */
typedef struct {
    int64_t pad_xy, n_vac_z, v_cycles, v_cycles_max;
    int32_t has_tol_abs, has_tol_rel;
    double tol_abs, tol_rel;
    int64_t pre_smooth, post_smooth;
    int32_t smoother;       /* 0 weighted Jacobi, 1 red-black SOR (MGSmoother :97-112) */
    int32_t warm_start;
    double omega, sor_omega;
    int32_t bc;             /* 0 Dirichlet zero, 1 dipole, 2 treecode (BoundaryCondition :81-94) */
    int32_t stencil;        /* 0 seven-point, 1 iso9+z, 2 iso27 (LaplacianStencilKind :708-712) */
    double tree_theta;
    int64_t tree_leaf, tree_max_depth;
    int32_t prolong;        /* 0 injection, 1 trilinear (ProlongationKind :714-718) */
    int32_t coarse_op;      /* 0 rediscretize, 1 Galerkin (CoarseOpKind :720-724) */
    double iso27_alpha;     /* <= 0 or non-finite: no override (iso27_alpha_bits == 0) */
} orc_mg_cfg;

#define ORC_MG_MAX_ST 729   /* Galerkin probe box is 9^3 */
#define ORC_MG_MAX_LEVELS 33

/* Stencil3D (demag_poisson_mg.rs:810-816) */
typedef struct {
    double center, diag;
    int32_t n;
    int32_t offs[ORC_MG_MAX_ST][3];
    double coeffs[ORC_MG_MAX_ST];
} orc_stencil;

typedef struct {
    int64_t cycles;          /* V-cycles run by the last solve */
    double rhs_max;          /* max |rhs| over interior cells (0 unless tol_rel is set) */
    double tol_target;
    double last_residual;    /* max |r| at the last check (-1 if none) */
    int32_t diverged;        /* NaN/Inf guard fired (phi reset to 0) */
    int32_t n_charges;       /* treecode charges of the last boundary update */
} orc_mg_stats;

typedef struct orc_mg orc_mg;

void orc_mg_cfg_default(orc_mg_cfg *c);                                  /* Default :166-185 */
void orc_mg_sanitize(orc_mg_cfg *c);                                     /* sanitize_cfg_for_op :1400-1417 */

/* stencils (demag_poisson_mg.rs:840-1189) */
void orc_stencil_seven_point(double dx, double dy, double dz, orc_stencil *s);   /* :840-872 */
void orc_stencil_iso9_plus_z(double dx, double dy, double dz, orc_stencil *s);   /* :874-922 */
void orc_stencil_iso27(double dx, double dy, double dz, double alpha, orc_stencil *s); /* :924-1058 */
void orc_stencil_galerkin(const orc_stencil *fine, int rx, int ry, int rz, int trilinear,
                          orc_stencil *out);                                      /* :1087-1189 */

/* kernels (mg_kernels.rs) */
void orc_mg_stamp_bc(double *arr, const double *bc, int64_t nx, int64_t ny, int64_t nz);   /* :35-61 */
void orc_mg_apply_stencil(const orc_stencil *s, const double *phi, double *out,
                          int64_t nx, int64_t ny, int64_t nz);                  /* apply_stencil_to_field :1193-1211 */
void orc_mg_smooth_wj(double *phi, double *tmp, const double *rhs, const double *bc,
                      int64_t nx, int64_t ny, int64_t nz, const orc_stencil *s,
                      int64_t iters, double omega);                             /* :121-162 */
void orc_mg_smooth_rb(double *phi, double *tmp, const double *rhs, const double *bc,
                      int64_t nx, int64_t ny, int64_t nz, double inv_dx2, double inv_dy2,
                      double inv_dz2, int64_t iters, double omega);             /* :172-240 */
double orc_mg_residual(const double *phi, const double *rhs, double *res, int64_t nx,
                       int64_t ny, int64_t nz, const orc_stencil *s);          /* :249-285 */
void orc_mg_restrict(const double *fine_res, int64_t fnx, int64_t fny, int64_t fnz,
                     double *coarse_rhs, double *coarse_phi, int64_t cnx, int64_t cny,
                     int64_t cnz);                                              /* :293-337 */
void orc_mg_prolong_add(const double *coarse, int64_t cnx, int64_t cny, int64_t cnz,
                        double *fine, int64_t fnx, int64_t fny, int64_t fnz, int trilinear); /* :349-439 */

/* treecode / dipole boundary values (mg_treecode.rs) */
int64_t orc_mg_build_charges(const double *rhs, int64_t nx, int64_t ny, int64_t nz, double dx,
                             double dy, double dz, double *pos_q /* 4 per charge, or NULL */); /* :273-300 */
void orc_mg_treecode_eval(const double *pos_q, int64_t n_charges, const double root_center[3],
                          double root_half, int64_t leaf, double theta, int64_t max_depth,
                          const double *targets, int64_t n_targets, double *phi_out);     /* :71-260 */
void orc_mg_treecode_bc(double *bc, const double *rhs, int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, int64_t leaf, double theta,
                        int64_t max_depth);                                    /* :308-366 */
void orc_mg_dipole_bc(double *bc, const double *rhs, int64_t nx, int64_t ny, int64_t nz,
                      double dx, double dy, double dz);                        /* :371-448 */

/* solver (DemagPoissonMG) */
orc_mg *orc_mg_new(const orc_grid *g, const orc_mg_cfg *cfg);            /* new_with_operator :1425-1577 */
void orc_mg_free(orc_mg *s);
int orc_mg_same_structure(const orc_mg *s, const orc_grid *g, const orc_mg_cfg *cfg); /* :1585-1593 */
void orc_mg_apply_cfg(orc_mg *s, const orc_mg_cfg *cfg);                  /* :1579-1583 */
void orc_mg_dims(const orc_mg *s, int64_t dims[6]);                       /* px py pz offx offy offz */
int orc_mg_levels(const orc_mg *s, int64_t *dims /* 3 per level */, double *spacing /* 3 per level */);
void orc_mg_level_stencil(const orc_mg *s, int level, orc_stencil *out);
void orc_mg_build_rhs(orc_mg *s, const double *m, double ms);             /* :1595-1802 */
void orc_mg_solve(orc_mg *s);                                             /* solve_with_timing :1947-2057 */
void orc_mg_add_b(const orc_mg *s, const double *m, double *b);           /* :2063-2092, mg_kernels.rs:450-507 */
void orc_mg_add_b_opt(const orc_mg *s, const double *m, double *b);       /* m NULL: every film cell (:2072-2074) */
int64_t orc_gaussian_kernel_1d(double sigma, double *w, int64_t cap);     /* mg_kernels.rs:580-599 */
void orc_mg_screen_rhs(orc_mg *s, double sigma);                          /* mg_kernels.rs:604-680 */
void orc_mg_add_field(orc_mg *s, const double *m, double *b, double ms);  /* :2094-2098 */
double *orc_mg_finest(orc_mg *s, int which);                              /* 0 phi, 1 rhs, 2 res, 3 bc */
void orc_mg_stats_get(const orc_mg *s, orc_mg_stats *out);

/* context hooks: the MG configuration a context's demag uses (default
 * DemagPoissonMGConfig::default() + default operator) and its solver */
void orc_ctx_set_mg_cfg(orc_ctx *c, const orc_mg_cfg *cfg);
orc_mg *orc_ctx_mg(orc_ctx *c);

#ifdef __cplusplus
}
#endif
#endif

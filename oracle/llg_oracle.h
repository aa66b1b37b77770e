/*
 * llg_oracle.h -- CPU FP64 restatement of llg-sim's uniform-grid LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke()) and the CPU baseline that
 * bench.py times beside the GPU.  Nothing in the product library links, loads
 * or calls it.
 *
 * Every function restates one reference function; the citation is given next
 * to each declaration (paths relative to the llg-sim source tree).
 *
 * Layouts follow the reference: vector fields are AoS doubles [cell][3] with
 * cell index j*nx + i (grid.rs:54-58, vector_field.rs:24-27); geometry masks
 * are one byte per cell (nonzero = material).
 *
 * Parity status per function (what pins it) is listed in oracle/README.md.
 */
#ifndef LLG_ORACLE_H
#define LLG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t nx, ny;
    double dx, dy, dz;
} orc_grid;                                   /* grid.rs:5-31 */

typedef struct {
    double ms, a_ex, k_u;
    double easy_axis[3];
    int32_t has_dmi;                          /* Option<f64>: 0 = None */
    double dmi;
    int32_t demag;
    int32_t demag_method;                     /* 0 fft, 1 mg, 2 dst */
} orc_material;                               /* params.rs:94-109 */

typedef struct {
    double gamma, alpha, dt;
    double b_ext[3];
} orc_params;                                 /* params.rs:137-144 */

enum { ORC_EXCH_ANIS = 0, ORC_EXCH_ANIS_DMI = 1, ORC_FULL = 2 };   /* mod.rs:37-45 */

#define ORC_MU0 (4.0 * 3.14159265358979323846 * 1e-7)               /* params.rs:15 */
#define ORC_GAMMA_E 1.76085963023e11                                 /* params.rs:16 */
#define ORC_DEMAG_ACCURACY 10.0                                      /* demag_fft_uniform.rs:432 */

/* ---- local field terms ------------------------------------------------ */
void orc_add_zeeman(int64_t n, double *b, const double b_ext[3]);             /* zeeman.rs:6-11 */
void orc_add_anisotropy(int64_t n, const double *m, double *b,
                        const orc_material *mat);                            /* anisotropy.rs:10-25 */
void orc_add_exchange(const orc_grid *g, const double *m, double *b,
                      const orc_material *mat, const uint8_t *mask);         /* exchange.rs:45-222 */
void orc_add_dmi(const orc_grid *g, const double *m, double *b,
                 const orc_material *mat, const uint8_t *mask);              /* dmi.rs:67-209 */
void orc_llg_rhs(int64_t n, const double *m, const double *b, double gamma,
                 double alpha, double *k);                                   /* llg.rs:20-32 */

/* ---- FFT (the rustfft boundary, restated from the DFT definition) ------- */
/* In place, unnormalised, sign -1 forward / +1 inverse, any length. */
void orc_fft1d(double *data_cplx, int64_t n, int inverse);
/* Row FFTs then column FFTs; inverse scaled 1/(nx*ny) (demag_fft_uniform.rs:777-922). */
void orc_fft2(double *data_cplx, int64_t nx, int64_t ny, int inverse);

/* ---- demag kernel (MuMax-like face-charge quadrature) ------------------- */
/* (mu0 Nxx, mu0 Nxy, mu0 Nyy, mu0 Nzz) for integer displacement (sx, sy).
   demag_fft_uniform.rs:1076-1097, 1139-1253 */
/* 0 = the reference's quadrature; 1, 2 = planted mistakes (tests only) */
void orc_set_quad_variant(int v);
void orc_kernel_components(double dx, double dy, double dz, int64_t sx, int64_t sy,
                           double accuracy, double out4[4]);
/* Row h = H from a unit M along source_axis (demag_fft_uniform.rs:1139-1242). */
void orc_h_from_unit_m(int source_axis, const double r_center[3], const double cell[3],
                       const int64_t disp[3], double accuracy, double h[3]);
/* Real-space padded kernels, px*py each, parity-reflected (demag_fft_uniform.rs:1013-1072). */
void orc_build_kernel_realspace(const orc_grid *g, int64_t px, int64_t py,
                                double *kxx, double *kxy, double *kyy, double *kzz);

/* Periodic real-space kernels (demag_fft_uniform.rs:941-1005): px = nx on a
   periodic axis (pbc > 0), 2n on an open one. */
void orc_build_kernel_realspace_pbc(const orc_grid *g, int64_t px, int64_t py, int64_t pbc_x,
                                    int64_t pbc_y, double *kxx, double *kxy, double *kyy,
                                    double *kzz);

/* ---- demag operator ------------------------------------------------------ */
typedef struct orc_demag orc_demag;
orc_demag *orc_demag_new(const orc_grid *g);                                 /* Demag2D::new :470-634 (open BC) */
orc_demag *orc_demag_new_pbc(const orc_grid *g, int64_t pbc_x, int64_t pbc_y);  /* Demag2D::new, any BC */
void orc_demag_free(orc_demag *d);
void orc_demag_add_field(orc_demag *d, const double *m, double *b, double ms);  /* :636-769 */
/* Access to the k-space kernels (complex, px*py each) for tests. */
void orc_demag_dims(const orc_demag *d, int64_t *px, int64_t *py);
void orc_demag_kspace(const orc_demag *d, double *kxx, double *kxy, double *kyy, double *kzz);

/* ---- context: field builder + RK45 ------------------------------------------- */
typedef struct orc_ctx orc_ctx;
orc_ctx *orc_ctx_new(const orc_grid *g);
void orc_ctx_free(orc_ctx *c);
/* build_h_eff_masked (mod.rs:48-75) when geom == NULL and use_geom == 0;
   build_h_eff_masked_geom (mod.rs:78-101) when use_geom != 0. */
void orc_build_h_eff(orc_ctx *c, const double *m, double *b, const orc_params *p,
                     const orc_material *mat, int field_mask, int use_geom,
                     const uint8_t *geom);
/* demag::add_demag_field (demag.rs:47-67) through the cached FFT operator. */
void orc_add_demag(orc_ctx *c, const double *m, double *b, const orc_material *mat);
/* demag::add_demag_field_pbc (demag.rs:70-143). */
void orc_add_demag_pbc(orc_ctx *c, const double *m, double *b, const orc_material *mat,
                       int64_t pbc_x, int64_t pbc_y);
/* step_llg_rk45_recompute_field_masked_adaptive (llg.rs:795-1026). */
void orc_rk45_step(orc_ctx *c, double *m, orc_params *p, const orc_material *mat,
                   int field_mask, double max_err, double headroom, double dt_min,
                   double dt_max, double *eps, int *accepted, double *dt_used);
void orc_rk45_reset(orc_ctx *c);
/* The context's grid. */
void orc_ctx_grid(const orc_ctx *c, orc_grid *g);
/* Number of field builds performed by this context (for FSAL accounting). */
int64_t orc_ctx_field_builds(const orc_ctx *c);

/* ---- threads ---- */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif

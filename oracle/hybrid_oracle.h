/*
 * hybrid_oracle.h -- CPU FP64 restatement of llg-sim's hybrid MG demag
 * (MG + delta-kernel correction).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef HYBRID_ORACLE_H
#define HYBRID_ORACLE_H

#include "llg_oracle.h"
#include "mg_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HybridConfig (demag_poisson_mg.rs:274-361) plus the boundary-scaling
   environment knobs LLG_DEMAG_MG_HYBRID_BSCALE / _BMARGIN (:2743-2760). */
typedef struct {
    int32_t enabled;
    int32_t reserved;
    int64_t radius_xy;
    int64_t delta_v_cycles;
    double sigma_cells;
    double bscale;
    int64_t bmargin;
} orc_hyb_cfg;

void orc_hyb_cfg_default(orc_hyb_cfg *h);
int orc_hyb_enabled(const orc_hyb_cfg *h);
int64_t orc_hyb_effective_radius(const orc_grid *g, const orc_hyb_cfg *h);
/* build_delta_kernel_impulse (:3159-3597): dk = [xx | xy | yy | zz], (2r+1)^2 each,
   index (dy + r)(2r+1) + (dx + r). */
void orc_hyb_delta_kernel(const orc_grid *g, const orc_mg_cfg *cfg, const orc_hyb_cfg *h, int64_t r,
                          double ms, double *dk);
/* DeltaKernel2D::add_correction (:410-583), accumulate into b */
void orc_hyb_add_correction(const orc_grid *g, int64_t r, const double *dk, const double *m, double *b,
                            double ms, double bscale, int64_t bmargin);
/* DemagPoissonMGHybrid::add_field, hybrid branch (:2515-2528) */
void orc_hyb_add_field(orc_mg *mg, const orc_hyb_cfg *h, int64_t r, const double *dk, const double *m,
                       double *b, double ms, const orc_grid *g);
/* the hybrid configuration a context's MG dispatch uses (default: disabled) */
void orc_ctx_set_hyb_cfg(orc_ctx *c, const orc_hyb_cfg *h);

#ifdef __cplusplus
}
#endif
#endif

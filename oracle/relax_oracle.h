/*
 * relax_oracle.h -- CPU FP64 restatement of llg-sim's relaxation path:
 * the damping-only Bogacki-Shampine RK23 stepper, the energy breakdown and the
 * MuMax-like relax controller.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path and the CPU baseline
 * of bench.py); the product library neither links nor loads it.
 *
 * Layouts as in llg_oracle.h (AoS doubles [cell][3], index j*nx + i).
 */
#ifndef RELAX_ORACLE_H
#define RELAX_ORACLE_H

#include "llg_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* llg_rhs_relax (llg.rs:38-50): k = -gamma*alpha/(1+alpha^2) m x (m x B) */
void orc_llg_rhs_relax(int64_t n, const double *m, const double *b, double gamma,
                       double alpha, double *k);

/* RK23Scratch (llg.rs:196-268): stage states, fields, torques, FSAL state. */
typedef struct orc_rk23 orc_rk23;
orc_rk23 *orc_rk23_new(const orc_grid *g);
void orc_rk23_free(orc_rk23 *s);
void orc_rk23_invalidate(orc_rk23 *s);                       /* :264-268 */
/* b4 when it belongs to the accepted state, else NULL (:252-258). */
const double *orc_rk23_last_b_eff(const orc_rk23 *s);

/* step_llg_rk23_recompute_field_masked_relax_adaptive (llg.rs:278-436). */
void orc_rk23_step(orc_ctx *c, orc_rk23 *s, double *m, orc_params *p,
                   const orc_material *mat, int field_mask, double max_err,
                   double headroom, double dt_min, double dt_max, double *eps,
                   int *accepted, double *dt_used);

/* compute_energy_geom (energy.rs:78-283): out = {exchange, anisotropy,
   zeeman, dmi, demag} in joules; geom may be NULL. */
void orc_energy(orc_ctx *c, const double *m, const orc_material *mat,
                const double b_ext[3], const uint8_t *geom, double out[5]);

enum { ORC_TQ_MAX = 0, ORC_TQ_MEAN = 1, ORC_TQ_RMS = 2, ORC_TQ_MEANSQ = 3 };   /* relax.rs:32-38 */
/* torque_metric_from_field (relax.rs:144-190) */
double orc_torque_metric(int64_t n, const double *m, const double *b, int metric);

typedef struct {
    double max_err, headroom, dt_min, dt_max;
    int32_t phase1_enabled, phase2_enabled;
    int64_t energy_stride;
    double rel_energy_tol;
    int32_t torque_metric;
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
} orc_relax_settings;                                        /* relax.rs:73-105 */

enum { ORC_STOP_MAX_ACCEPTED = 0, ORC_STOP_PHASE2_DISABLED = 1,
       ORC_STOP_TIGHTEN_FLOOR = 2, ORC_STOP_GATE_NOT_SATISFIED = 3 };    /* relax.rs:40-46 */
enum { ORC_STAGE_NONE = -1, ORC_STAGE_BELOW_THRESHOLD = 0, ORC_STAGE_PLATEAU = 1,
       ORC_STAGE_MAX_ACCEPTED = 2 };                                     /* relax.rs:48-53 */

typedef struct {
    int64_t accepted_steps, rejected_steps, torque_checks, torque_field_rebuilds;
    int32_t stop_reason, last_stage_stop;
    int32_t has_final_torque, has_final_torque_max;
    double final_torque, final_torque_max, final_max_err, final_dt;
} orc_relax_report;                                          /* relax.rs:55-71 */

void orc_relax_settings_default(orc_relax_settings *s);      /* relax.rs:107-139 */
void orc_relax_settings_mumax_like(orc_relax_settings *s);   /* relax.rs:789-808 */
/* relax_with_report (relax.rs:334-774); settings->max_err is written back. */
void orc_relax(orc_ctx *c, orc_rk23 *s, double *m, orc_params *p, const orc_material *mat,
               int field_mask, orc_relax_settings *settings, orc_relax_report *rep);

#ifdef __cplusplus
}
#endif
#endif

/*
 * relax_oracle.c -- CPU FP64 restatement of llg-sim's relaxation path
 * (llg.rs:38-50, 196-436; energy.rs:36-283; relax.rs:32-808).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path and the CPU baseline
 * of bench.py.  Loops, accumulation order and branch structure follow the
 * reference so that results agree to rounding.
 */
#include "relax_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
/* normalize(0) = (0,0,1) (vec3.rs:21-28) */
static inline void normalize3(const double *v, double *o) {
    double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (n2 == 0.0) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 1.0;
        return;
    }
    double inv = 1.0 / sqrt(n2);
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

/* llg.rs:38-50 */
void orc_llg_rhs_relax(int64_t n, const double *m, const double *b, double gamma,
                       double alpha, double *k) {
    double denom = 1.0 + alpha * alpha;
    double pref = -gamma * alpha / denom;
    for (int64_t i = 0; i < n; ++i) {
        double mxb[3], mxmxb[3];
        cross3(m + 3 * i, b + 3 * i, mxb);
        cross3(m + 3 * i, mxb, mxmxb);
        k[3 * i] = pref * mxmxb[0];
        k[3 * i + 1] = pref * mxmxb[1];
        k[3 * i + 2] = pref * mxmxb[2];
    }
}

/* ------------------------------------------------------------------------ */
/* RK23Scratch and the relax-mode Bogacki-Shampine step                      */
/* ------------------------------------------------------------------------ */
struct orc_rk23 {
    int64_t n;
    double *m2, *m3, *m4, *b1, *b2, *b3, *b4, *k1, *k2, *k3, *k4;
    int last_b_eff_valid;
    int has_last;          /* last_mask / last_b_ext are Some */
    int last_mask;
    double last_b_ext[3];
};

orc_rk23 *orc_rk23_new(const orc_grid *g) {
    orc_rk23 *s = (orc_rk23 *)calloc(1, sizeof(orc_rk23));
    s->n = g->nx * g->ny;
    size_t n3 = (size_t)(3 * s->n);
    double **bufs[] = {&s->m2, &s->m3, &s->m4, &s->b1, &s->b2, &s->b3, &s->b4,
                       &s->k1, &s->k2, &s->k3, &s->k4};
    for (size_t q = 0; q < sizeof(bufs) / sizeof(bufs[0]); ++q)
        *bufs[q] = (double *)calloc(n3 ? n3 : 1, sizeof(double));
    return s;
}

void orc_rk23_free(orc_rk23 *s) {
    if (!s) return;
    double *bufs[] = {s->m2, s->m3, s->m4, s->b1, s->b2, s->b3, s->b4, s->k1, s->k2, s->k3, s->k4};
    for (size_t q = 0; q < sizeof(bufs) / sizeof(bufs[0]); ++q) free(bufs[q]);
    free(s);
}

void orc_rk23_invalidate(orc_rk23 *s) {
    s->last_b_eff_valid = 0;
    s->has_last = 0;
}

const double *orc_rk23_last_b_eff(const orc_rk23 *s) {
    return s->last_b_eff_valid ? s->b4 : NULL;
}

/* m_out = normalize(m + (dt*c_0) k_0 + (dt*c_1) k_1 + ...), added in order */
static void stage_combo(int64_t n, const double *m, int nk, const double *const *ks,
                        const double *cs, double dt0, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        double v[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
        for (int q = 0; q < nk; ++q) {
            double sc = dt0 * cs[q];
            const double *k = ks[q] + 3 * i;
            v[0] = v[0] + sc * k[0];
            v[1] = v[1] + sc * k[1];
            v[2] = v[2] + sc * k[2];
        }
        normalize3(v, out + 3 * i);
    }
}

/* step_llg_rk23_recompute_field_masked_relax_adaptive (llg.rs:278-436) */
void orc_rk23_step(orc_ctx *c, orc_rk23 *s, double *m, orc_params *p,
                   const orc_material *mat, int field_mask, double max_err,
                   double headroom, double dt_min, double dt_max, double *eps_out,
                   int *accepted_out, double *dt_used_out) {
    const double A21 = 1.0 / 2.0, A32 = 3.0 / 4.0;
    const double B1 = 2.0 / 9.0, B2 = 1.0 / 3.0, B3 = 4.0 / 9.0;
    const double BH1 = 7.0 / 24.0, BH2 = 1.0 / 4.0, BH3 = 1.0 / 3.0, BH4 = 1.0 / 8.0;
    const int64_t n = s->n;
    const size_t bytes = sizeof(double) * 3 * (size_t)n;
    const double gamma = p->gamma, alpha = p->alpha, dt0 = p->dt;

    int fsal_ok = s->last_b_eff_valid && s->has_last && s->last_mask == field_mask &&
                  s->last_b_ext[0] == p->b_ext[0] && s->last_b_ext[1] == p->b_ext[1] &&
                  s->last_b_ext[2] == p->b_ext[2];
    if (fsal_ok) {
        memcpy(s->b1, s->b4, bytes);
        memcpy(s->k1, s->k4, bytes);
    } else {
        orc_build_h_eff(c, m, s->b1, p, mat, field_mask, 0, NULL);
        orc_llg_rhs_relax(n, m, s->b1, gamma, alpha, s->k1);
    }

    {   /* stage 2 */
        const double *ks[1] = {s->k1};
        const double cs[1] = {A21};
        stage_combo(n, m, 1, ks, cs, dt0, s->m2);
        orc_build_h_eff(c, s->m2, s->b2, p, mat, field_mask, 0, NULL);
        orc_llg_rhs_relax(n, s->m2, s->b2, gamma, alpha, s->k2);
    }
    {   /* stage 3 */
        const double *ks[1] = {s->k2};
        const double cs[1] = {A32};
        stage_combo(n, m, 1, ks, cs, dt0, s->m3);
        orc_build_h_eff(c, s->m3, s->b3, p, mat, field_mask, 0, NULL);
        orc_llg_rhs_relax(n, s->m3, s->b3, gamma, alpha, s->k3);
    }
    {   /* third-order solution and its derivative */
        const double *ks[3] = {s->k1, s->k2, s->k3};
        const double cs[3] = {B1, B2, B3};
        stage_combo(n, m, 3, ks, cs, dt0, s->m4);
        orc_build_h_eff(c, s->m4, s->b4, p, mat, field_mask, 0, NULL);
        orc_llg_rhs_relax(n, s->m4, s->b4, gamma, alpha, s->k4);
    }
    {   /* embedded second-order solution (into m2) */
        const double *ks[4] = {s->k1, s->k2, s->k3, s->k4};
        const double cs[4] = {BH1, BH2, BH3, BH4};
        stage_combo(n, m, 4, ks, cs, dt0, s->m2);
    }
    double err_inf = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double *a = s->m4 + 3 * i, *b = s->m2 + 3 * i;
        double d = fmax(fmax(fabs(a[0] - b[0]), fabs(a[1] - b[1])), fabs(a[2] - b[2]));
        if (d > err_inf) err_inf = d;
    }
    const double eps = err_inf;

    double dt_next = eps > 0.0 ? dt0 * headroom * pow(max_err / eps, 1.0 / 3.0) : dt0 * 2.0;
    if (dt_next < dt_min) dt_next = dt_min;
    if (dt_next > dt_max) dt_next = dt_max;
    int accept = (eps <= max_err) || (dt0 <= dt_min * 1.0000000001);
    if (accept) {
        memcpy(m, s->m4, bytes);
        s->last_b_eff_valid = 1;
        s->has_last = 1;
        s->last_mask = field_mask;
        s->last_b_ext[0] = p->b_ext[0]; s->last_b_ext[1] = p->b_ext[1]; s->last_b_ext[2] = p->b_ext[2];
    } else {
        s->last_b_eff_valid = 0;
        s->has_last = 0;
    }
    p->dt = dt_next;
    *eps_out = eps;
    *accepted_out = accept;
    *dt_used_out = dt0;
}

/* ------------------------------------------------------------------------ */
/* Energy breakdown (energy.rs:36-283)                                        */
/* ------------------------------------------------------------------------ */
static inline void e_ghost_x(const double *m, double n_x, double eta, double dx, double *o) {
    o[0] = m[0] + n_x * dx * (-eta * m[2]);
    o[1] = m[1] + n_x * dx * 0.0;
    o[2] = m[2] + n_x * dx * (eta * m[0]);
}
static inline void e_ghost_y(const double *m, double n_y, double eta, double dy, double *o) {
    o[0] = m[0] + n_y * dy * 0.0;
    o[1] = m[1] + n_y * dy * (-eta * m[2]);
    o[2] = m[2] + n_y * dy * (eta * m[1]);
}

void orc_energy(orc_ctx *c, const double *m, const orc_material *mat,
                const double b_ext[3], const uint8_t *geom, double out[5]) {
    orc_grid g;
    orc_ctx_grid(c, &g);
    const int64_t nx = g.nx, ny = g.ny;
    const double dx = g.dx, dy = g.dy;
    const double v = g.dx * g.dy * g.dz;   /* Grid2D::cell_volume (grid.rs:25-27) */
    const int64_t n = nx * ny;
    const double aex = mat->a_ex, ku = mat->k_u, ms = mat->ms;
    const double *u = mat->easy_axis;
    const double bx = b_ext[0], by = b_ext[1], bz = b_ext[2];
    double e_ex = 0.0, e_an = 0.0, e_zee = 0.0, e_dmi = 0.0, e_demag = 0.0;

    const int dmi_on = mat->has_dmi && mat->dmi != 0.0 && mat->a_ex != 0.0;
    const double d = mat->dmi;
    const double eta = dmi_on ? d / (2.0 * aex) : 0.0;

    double *bd = NULL;
    if (mat->demag) {
        bd = (double *)calloc((size_t)(3 * n), sizeof(double));
        orc_add_demag(c, m, bd, mat);
    }
#define INSIDE(q) (geom == NULL || geom[(q)] != 0)
    for (int64_t j = 0; j < ny; ++j) {
        for (int64_t i = 0; i < nx; ++i) {
            const int64_t idx = j * nx + i;
            if (!INSIDE(idx)) continue;
            const double *mij = m + 3 * idx;
            const double mx = mij[0], my = mij[1], mz = mij[2];

            if (aex != 0.0) {
                if (i + 1 < nx) {
                    const int64_t r = idx + 1;
                    if (INSIDE(r)) {
                        const double *mp = m + 3 * r;
                        double a0 = mp[0] - mx, a1 = mp[1] - my, a2 = mp[2] - mz;
                        double sq = a0 * a0 + a1 * a1 + a2 * a2;
                        e_ex += aex * (sq / (dx * dx)) * v;
                    }
                }
                if (j + 1 < ny) {
                    const int64_t up = idx + nx;
                    if (INSIDE(up)) {
                        const double *mp = m + 3 * up;
                        double a0 = mp[0] - mx, a1 = mp[1] - my, a2 = mp[2] - mz;
                        double sq = a0 * a0 + a1 * a1 + a2 * a2;
                        e_ex += aex * (sq / (dy * dy)) * v;
                    }
                }
            }
            if (ku != 0.0) {
                double mdotu = mx * u[0] + my * u[1] + mz * u[2];
                e_an += ku * (1.0 - mdotu * mdotu) * v;
            }
            if (ms != 0.0) {
                double mdotb = mx * bx + my * by + mz * bz;
                e_zee -= ms * mdotb * v;
            }
            if (dmi_on && ms != 0.0) {
                const double pref = 2.0 * d / ms;
                double im[3], ip[3], jm[3], jp[3];
                if (nx == 1) {
                    memcpy(im, mij, sizeof im); memcpy(ip, mij, sizeof ip);
                } else {
                    if (i == 0) e_ghost_x(mij, -1.0, eta, dx, im);
                    else if (INSIDE(idx - 1)) memcpy(im, m + 3 * (idx - 1), sizeof im);
                    else e_ghost_x(mij, -1.0, eta, dx, im);
                    if (i == nx - 1) e_ghost_x(mij, 1.0, eta, dx, ip);
                    else if (INSIDE(idx + 1)) memcpy(ip, m + 3 * (idx + 1), sizeof ip);
                    else e_ghost_x(mij, 1.0, eta, dx, ip);
                }
                if (ny == 1) {
                    memcpy(jm, mij, sizeof jm); memcpy(jp, mij, sizeof jp);
                } else {
                    if (j == 0) e_ghost_y(mij, -1.0, eta, dy, jm);
                    else if (INSIDE(idx - nx)) memcpy(jm, m + 3 * (idx - nx), sizeof jm);
                    else e_ghost_y(mij, -1.0, eta, dy, jm);
                    if (j == ny - 1) e_ghost_y(mij, 1.0, eta, dy, jp);
                    else if (INSIDE(idx + nx)) memcpy(jp, m + 3 * (idx + nx), sizeof jp);
                    else e_ghost_y(mij, 1.0, eta, dy, jp);
                }
                double dmz_dx = nx > 1 ? (ip[2] - im[2]) / (2.0 * dx) : 0.0;
                double dmz_dy = ny > 1 ? (jp[2] - jm[2]) / (2.0 * dy) : 0.0;
                double dmx_dx = nx > 1 ? (ip[0] - im[0]) / (2.0 * dx) : 0.0;
                double dmy_dy = ny > 1 ? (jp[1] - jm[1]) / (2.0 * dy) : 0.0;
                double bdx = pref * dmz_dx, bdy = pref * dmz_dy, bdz = -pref * (dmx_dx + dmy_dy);
                double mdotb = mx * bdx + my * bdy + mz * bdz;
                e_dmi += -0.5 * ms * mdotb * v;
            }
            if (mat->demag && ms != 0.0) {
                const double *b = bd + 3 * idx;
                double mdotb = mx * b[0] + my * b[1] + mz * b[2];
                e_demag += -0.5 * ms * mdotb * v;
            }
        }
    }
#undef INSIDE
    free(bd);
    out[0] = e_ex; out[1] = e_an; out[2] = e_zee; out[3] = e_dmi; out[4] = e_demag;
}

static double energy_total(orc_ctx *c, const double *m, const orc_material *mat, const double b_ext[3]) {
    double e[5];
    orc_energy(c, m, mat, b_ext, NULL, e);
    return e[0] + e[1] + e[2] + e[3] + e[4];   /* EnergyBreakdown::total (energy.rs:30-34) */
}

/* ------------------------------------------------------------------------ */
/* Relax controller (relax.rs:144-808)                                       */
/* ------------------------------------------------------------------------ */
double orc_torque_metric(int64_t n, const double *m, const double *b, int metric) {
    double nf = (double)n;
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double t[3];
        cross3(m + 3 * i, b + 3 * i, t);
        double mag2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        if (metric == ORC_TQ_MAX) {
            double mag = sqrt(mag2);
            if (mag > acc) acc = mag;
        } else if (metric == ORC_TQ_MEAN) {
            acc += sqrt(mag2);
        } else {
            acc += mag2;
        }
    }
    switch (metric) {
        case ORC_TQ_MAX: return acc;
        case ORC_TQ_MEAN: return acc / fmax(nf, 1.0);
        case ORC_TQ_RMS: return sqrt(acc / fmax(nf, 1.0));
        default: return acc / fmax(nf, 1.0);
    }
}

void orc_relax_settings_default(orc_relax_settings *s) {
    memset(s, 0, sizeof *s);
    s->max_err = 1e-5;
    s->headroom = 0.8;
    s->dt_min = 1e-18;
    s->dt_max = 1e-11;
    s->phase1_enabled = 1;
    s->phase2_enabled = 1;
    s->energy_stride = 3;
    s->rel_energy_tol = 1e-12;
    s->torque_metric = ORC_TQ_MAX;
    s->has_torque_threshold = 1;
    s->torque_threshold = 1e-4;
    s->torque_check_stride = 1;
    s->torque_plateau_rel = 1e-3;
    s->torque_plateau_abs = 0.0;
    s->torque_plateau_checks = 0;
    s->torque_plateau_min_checks = 5;
    s->tighten_factor = 0.70710678118654752440;   /* FRAC_1_SQRT_2 */
    s->tighten_floor = 1e-9;
    s->max_accepted_steps = 2000000;
}

void orc_relax_settings_mumax_like(orc_relax_settings *s) {
    orc_relax_settings_default(s);
    s->phase1_enabled = 1;
    s->phase2_enabled = 1;
    s->energy_stride = 3;
    s->rel_energy_tol = 0.0;
    s->torque_metric = ORC_TQ_MEANSQ;
    s->has_torque_threshold = 0;
    s->torque_threshold = 0.0;
    s->torque_plateau_checks = 1;
    s->torque_plateau_min_checks = 1;
    s->torque_plateau_rel = 0.0;
    s->torque_plateau_abs = 0.0;
    s->tighten_factor = 0.70710678118654752440;
    s->tighten_floor = 1e-9;
}

typedef struct {
    orc_ctx *c;
    orc_rk23 *s;
    double *m;
    orc_params *p;
    const orc_material *mat;
    int mask;
    double *b_scratch;
    int64_t torque_checks, torque_field_rebuilds, accepted, rejected;
} relax_run;

/* torque_now (relax.rs:206-230) */
static double torque_now(relax_run *r, int metric) {
    r->torque_checks += 1;
    const double *b = orc_rk23_last_b_eff(r->s);
    if (!b) {
        r->torque_field_rebuilds += 1;
        orc_build_h_eff(r->c, r->m, r->b_scratch, r->p, r->mat, r->mask, 0, NULL);
        b = r->b_scratch;
    }
    return orc_torque_metric(r->s->n, r->m, b, metric);
}

/* advance_accepted (relax.rs:247-277) */
static int advance_accepted(relax_run *r, double max_err, double headroom, double dt_min,
                            double dt_max, int64_t max_accepted, int64_t n_accept) {
    int64_t target = r->accepted + (n_accept > 1 ? n_accept : 1);
    while (r->accepted < target) {
        if (r->accepted >= max_accepted) return 0;
        double eps, dtu;
        int ok;
        orc_rk23_step(r->c, r->s, r->m, r->p, r->mat, r->mask, max_err, headroom, dt_min, dt_max,
                      &eps, &ok, &dtu);
        if (ok) r->accepted += 1;
        else r->rejected += 1;
    }
    return 1;
}

static void fill_report(orc_relax_report *rep, const relax_run *r, int stop, int stage,
                        int has_t, double t, int has_tmax, double tmax, double max_err, double dt) {
    rep->accepted_steps = r->accepted;
    rep->rejected_steps = r->rejected;
    rep->torque_checks = r->torque_checks;
    rep->torque_field_rebuilds = r->torque_field_rebuilds;
    rep->stop_reason = stop;
    rep->last_stage_stop = stage;
    rep->has_final_torque = has_t;
    rep->final_torque = has_t ? t : 0.0;
    rep->has_final_torque_max = has_tmax;
    rep->final_torque_max = has_tmax ? tmax : 0.0;
    rep->final_max_err = max_err;
    rep->final_dt = dt;
}

static inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

void orc_relax(orc_ctx *c, orc_rk23 *s, double *m, orc_params *p, const orc_material *mat,
               int field_mask, orc_relax_settings *st, orc_relax_report *rep) {
    const int phase1 = st->phase1_enabled, phase2 = st->phase2_enabled;
    const int64_t energy_stride = st->energy_stride > 1 ? st->energy_stride : 1;
    const double rel_energy_tol = st->rel_energy_tol;
    const int metric = st->torque_metric;
    const int has_thr = st->has_torque_threshold;
    const double thr = st->torque_threshold;
    const int64_t check_stride = st->torque_check_stride > 1 ? st->torque_check_stride : 1;
    const double plateau_rel = st->torque_plateau_rel, plateau_abs = st->torque_plateau_abs;
    const int64_t plateau_checks = st->torque_plateau_checks;
    const int64_t plateau_min_checks = st->torque_plateau_min_checks > 1 ? st->torque_plateau_min_checks : 1;
    const double tighten_factor = st->tighten_factor, tighten_floor = st->tighten_floor;
    const int64_t max_acc = st->max_accepted_steps;
    const int has_gate = st->has_final_torque_gate, has_gate_max = st->has_final_torque_gate_max;
    const double gate = st->final_torque_gate, gate_max = st->final_torque_gate_max;
    const int64_t gate_extra = st->gate_max_extra_accepted_steps;
    const int64_t gate_plateau_limit = st->gate_plateau_fails;
    double max_err = st->max_err;

    p->dt = clampd(p->dt, st->dt_min, st->dt_max);
    orc_rk23_invalidate(s);

    relax_run r = {c, s, m, p, mat, field_mask, NULL, 0, 0, 0, 0};
    r.b_scratch = (double *)calloc((size_t)(3 * s->n) + 1, sizeof(double));
    int last_stage = ORC_STAGE_NONE;
    int has_ft = 0, has_ftmax = 0;
    double ft = 0.0, ftmax = 0.0;
    const double headroom = st->headroom, dt_min = st->dt_min, dt_max = st->dt_max;

    if (phase1) {
        double e0 = energy_total(c, m, mat, p->b_ext);
        for (;;) {
            if (!advance_accepted(&r, max_err, headroom, dt_min, dt_max, max_acc, energy_stride)) {
                st->max_err = max_err;
                fill_report(rep, &r, ORC_STOP_MAX_ACCEPTED, last_stage, has_ft, ft, has_ftmax, ftmax,
                            max_err, p->dt);
                goto done;
            }
            double e1 = energy_total(c, m, mat, p->b_ext);
            double tol = rel_energy_tol * fmax(fabs(e0), 1e-30);
            if (e1 < e0 - tol) {
                e0 = e1;
                continue;
            }
            break;
        }
    }

    if (!phase2) {
        st->max_err = max_err;
        fill_report(rep, &r, ORC_STOP_PHASE2_DISABLED, last_stage, 0, 0.0, 0, 0.0, max_err, p->dt);
        goto done;
    }

    {
        const int plateau_enabled = !has_thr && plateau_checks > 0;
        const int64_t need_fails = plateau_checks > 1 ? plateau_checks : 1;
        for (;;) {
            double t_prev = torque_now(&r, metric);
            const int mumax_f32 = plateau_enabled && metric == ORC_TQ_MEANSQ && plateau_rel == 0.0 &&
                                  plateau_abs == 0.0 && plateau_checks == 1 && plateau_min_checks == 1;
            const double n_f = (double)s->n;
            float t_prev_f32 = mumax_f32 ? (float)(t_prev * n_f) : (float)t_prev;
            int64_t plateau_fails = 0, comparisons = 0;
            for (;;) {
                if (has_thr && t_prev <= thr) {
                    last_stage = ORC_STAGE_BELOW_THRESHOLD;
                    break;
                }
                if (plateau_enabled && plateau_fails >= need_fails) {
                    last_stage = ORC_STAGE_PLATEAU;
                    break;
                }
                if (r.accepted >= max_acc ||
                    !advance_accepted(&r, max_err, headroom, dt_min, dt_max, max_acc, check_stride)) {
                    st->max_err = max_err;
                    fill_report(rep, &r, ORC_STOP_MAX_ACCEPTED, ORC_STAGE_MAX_ACCEPTED, 1, t_prev,
                                has_ftmax, ftmax, max_err, p->dt);
                    goto done;
                }
                double t_new = torque_now(&r, metric);
                if (plateau_enabled) {
                    comparisons += 1;
                    if (comparisons < plateau_min_checks) {
                        plateau_fails = 0;
                    } else if (mumax_f32) {
                        float t_new_f32 = (float)(t_new * n_f);
                        int improved = t_new_f32 < t_prev_f32;
                        t_prev_f32 = t_new_f32;
                        plateau_fails = improved ? 0 : plateau_fails + 1;
                    } else {
                        double need = fmax(plateau_rel * fmax(fabs(t_prev), 1e-30), plateau_abs);
                        int improved = (t_prev - t_new) > need;
                        plateau_fails = improved ? 0 : plateau_fails + 1;
                    }
                }
                t_prev = t_new;
            }
            has_ft = 1;
            ft = t_prev;
            if (max_err <= tighten_floor) break;
            max_err *= tighten_factor;
            if (max_err < tighten_floor) max_err = tighten_floor;
        }
    }

    has_ftmax = 1;
    ftmax = torque_now(&r, ORC_TQ_MAX);

    if ((has_gate || has_gate_max) && gate_extra > 0) {
        const int64_t start = r.accepted;
        double last_t = has_ft ? ft : INFINITY;
        int64_t gate_fails = 0;
        for (;;) {
            /* gates_ok (relax.rs:279-332) */
            double t_metric = torque_now(&r, metric);
            const int need_max = has_gate_max || metric == ORC_TQ_MAX;
            double t_max = NAN;
            if (need_max) t_max = metric == ORC_TQ_MAX ? t_metric : torque_now(&r, ORC_TQ_MAX);
            int ok = (!has_gate || t_metric <= gate) && (!has_gate_max || t_max <= gate_max);
            has_ft = 1;
            ft = t_metric;
            if (isfinite(t_max)) { has_ftmax = 1; ftmax = t_max; }
            if (ok) break;
            int64_t used = r.accepted > start ? r.accepted - start : 0;
            if (used >= gate_extra) {
                st->max_err = max_err;
                fill_report(rep, &r, ORC_STOP_GATE_NOT_SATISFIED, last_stage, has_ft, ft, has_ftmax,
                            ftmax, max_err, p->dt);
                goto done;
            }
            if (gate_plateau_limit > 0) {
                double need = fmax(plateau_rel * fmax(fabs(last_t), 1e-30), plateau_abs);
                int improved = (last_t - t_metric) > need;
                gate_fails = improved ? 0 : gate_fails + 1;
                if (gate_fails >= gate_plateau_limit) {
                    st->max_err = max_err;
                    fill_report(rep, &r, ORC_STOP_GATE_NOT_SATISFIED, last_stage, has_ft, ft,
                                has_ftmax, ftmax, max_err, p->dt);
                    goto done;
                }
                last_t = t_metric;
            }
            if (!advance_accepted(&r, max_err, headroom, dt_min, dt_max, max_acc, check_stride)) {
                st->max_err = max_err;
                fill_report(rep, &r, ORC_STOP_MAX_ACCEPTED, ORC_STAGE_MAX_ACCEPTED, has_ft, ft,
                            has_ftmax, ftmax, max_err, p->dt);
                goto done;
            }
        }
    }

    st->max_err = max_err;
    fill_report(rep, &r, ORC_STOP_TIGHTEN_FLOOR, last_stage, has_ft, ft, has_ftmax, ftmax, max_err, p->dt);
done:
    free(r.b_scratch);
}

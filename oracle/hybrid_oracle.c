/*
 * hybrid_oracle.c -- CPU FP64 restatement of llg-sim's hybrid multigrid demag
 * (MG + local delta-kernel correction, demag_poisson_mg.rs:270-700, 2277-2530,
 * 3159-3597).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  The diagnostic
 * branches of the reference (LLG_DEMAG_MG_HYBRID_DIAG*, which only print) and
 * its on-disk delta-kernel cache (which only avoids rebuilds) are not restated.
 */
#include "hybrid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_hyb_cfg_default(orc_hyb_cfg *h) {   /* HybridConfig::default (:283-293) */
    memset(h, 0, sizeof *h);
    h->enabled = 0;
    h->radius_xy = 12;
    h->delta_v_cycles = 60;
    h->sigma_cells = 1.5;
    h->bscale = 0.0;
    h->bmargin = 2;
}

int orc_hyb_enabled(const orc_hyb_cfg *h) { return h->enabled && h->radius_xy > 0; }   /* :357-360 */

/* ensure_delta_kernel's effective radius (:2419-2425) */
int64_t orc_hyb_effective_radius(const orc_grid *g, const orc_hyb_cfg *h) {
    int64_t mx = (g->nx >= 2 ? g->nx - 2 : 0) / 2, my = (g->ny >= 2 ? g->ny - 2 : 0) / 2;
    int64_t r = h->radius_xy;
    if (r > mx) r = mx;
    if (r > my) r = my;
    return r;
}

/* DeltaKernel2D::symmetrize (:664-694): in place, in the reference's loop order */
static void symmetrize(int64_t r, double *xx, double *xy, double *yy, double *zz, double dx, double dy) {
    const int64_t st = 2 * r + 1;
#define KI(a, b) (((b) + r) * st + ((a) + r))
    for (int64_t b = -r; b <= r; ++b)
        for (int64_t a = -r; a <= r; ++a) {
            int64_t k = KI(a, b), ki = KI(-a, -b);
            xx[k] = 0.5 * (xx[k] + xx[ki]);
            yy[k] = 0.5 * (yy[k] + yy[ki]);
            zz[k] = 0.5 * (zz[k] + zz[ki]);
            xy[k] = 0.5 * (xy[k] + xy[ki]);
        }
    if (fabs(dx - dy) <= 1e-12 * fmax(fmax(fabs(dx), fabs(dy)), 1.0)) {
        for (int64_t b = -r; b <= r; ++b)
            for (int64_t a = -r; a <= r; ++a) {
                int64_t k = KI(a, b), ks = KI(b, a);
                double nxx = 0.5 * (xx[k] + yy[ks]);
                double nyy = 0.5 * (yy[k] + xx[ks]);
                double nxy = 0.5 * (xy[k] + xy[ks]);
                xx[k] = nxx;
                yy[k] = nyy;
                xy[k] = nxy;
            }
    }
#undef KI
}

/* build_delta_kernel_impulse (:3159-3597), the non-diagnostic part. */
void orc_hyb_delta_kernel(const orc_grid *g, const orc_mg_cfg *cfg, const orc_hyb_cfg *h, int64_t r,
                          double ms, double *dk) {
    const int64_t nx = g->nx, ny = g->ny, n = nx * ny;
    const int64_t cx = nx / 2, cy = ny / 2;
    const double ms_inv = fabs(ms) > 0.0 ? 1.0 / ms : 0.0;
    const double geom_eps = 1e-14;
    const int64_t st = 2 * r + 1, nst = st * st;
    double *xx = dk, *xy = dk + nst, *yy = dk + 2 * nst, *zz = dk + 3 * nst;
    memset(dk, 0, sizeof(double) * 4 * (size_t)nst);

    orc_mg_cfg ci = *cfg;
    ci.warm_start = 0;
    ci.has_tol_abs = 0;
    ci.has_tol_rel = 0;
    ci.v_cycles = h->delta_v_cycles;
    ci.v_cycles_max = h->delta_v_cycles > 1 ? h->delta_v_cycles : 1;
    orc_mg *mg = orc_mg_new(g, &ci);
    orc_demag *fft = orc_demag_new(g);

    double *m = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *bf = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *bm = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *xy_from_x = (double *)calloc((size_t)nst, sizeof(double));
    double *xy_from_y = (double *)calloc((size_t)nst, sizeof(double));

    for (int axis = 0; axis < 3; ++axis) {
        for (int64_t c = 0; c < n; ++c) { m[3 * c] = geom_eps; m[3 * c + 1] = 0.0; m[3 * c + 2] = 0.0; }
        const int64_t center = cy * nx + cx;
        if (axis == 0) { m[3 * center] = 1.0 + geom_eps; }
        else if (axis == 1) { m[3 * center + 1] = 1.0; }
        else { m[3 * center + 2] = 1.0; }
        memset(bf, 0, sizeof(double) * 3 * (size_t)n);
        orc_demag_add_field(fft, m, bf, ms);
        memset(bm, 0, sizeof(double) * 3 * (size_t)n);
        orc_mg_build_rhs(mg, m, ms);
        if (h->sigma_cells > 0.0) orc_mg_screen_rhs(mg, h->sigma_cells);
        orc_mg_solve(mg);
        orc_mg_add_b_opt(mg, NULL, bm);
        for (int64_t b = -r; b <= r; ++b)
            for (int64_t a = -r; a <= r; ++a) {
                const int64_t id = (cy + b) * nx + (cx + a);
                const int64_t k = (b + r) * st + (a + r);
                if (axis == 0) {
                    xx[k] = (bf[3 * id] - bm[3 * id]) * ms_inv;
                    xy_from_x[k] = (bf[3 * id + 1] - bm[3 * id + 1]) * ms_inv;
                } else if (axis == 1) {
                    yy[k] = (bf[3 * id + 1] - bm[3 * id + 1]) * ms_inv;
                    xy_from_y[k] = (bf[3 * id] - bm[3 * id]) * ms_inv;
                } else {
                    zz[k] = (bf[3 * id + 2] - bm[3 * id + 2]) * ms_inv;
                }
            }
    }
    for (int64_t k = 0; k < nst; ++k) xy[k] = 0.5 * (xy_from_x[k] + xy_from_y[k]);
    symmetrize(r, xx, xy, yy, zz, g->dx, g->dy);
    /* uniform-bias fix: the stencil sums to zero on the diagonal terms */
    const int64_t center = r * st + r;
    double sxx = 0.0, syy = 0.0, szz = 0.0;
    for (int64_t k = 0; k < nst; ++k) sxx += xx[k];
    for (int64_t k = 0; k < nst; ++k) syy += yy[k];
    for (int64_t k = 0; k < nst; ++k) szz += zz[k];
    xx[center] -= sxx;
    yy[center] -= syy;
    zz[center] -= szz;

    free(m); free(bf); free(bm); free(xy_from_x); free(xy_from_y);
    orc_mg_free(mg);
    orc_demag_free(fft);
}

/* DeltaKernel2D::add_correction (:410-583), with the boundary scaling of
 * LLG_DEMAG_MG_HYBRID_BSCALE > 0 (material fraction^alpha at cells within
 * `margin` of vacuum). */
void orc_hyb_add_correction(const orc_grid *g, int64_t r, const double *dk, const double *m, double *b,
                            double ms, double bscale, int64_t bmargin) {
    const int64_t nx = g->nx, ny = g->ny, n = nx * ny, st = 2 * r + 1, nst = st * st;
    const double *xx = dk, *xy = dk + nst, *yy = dk + 2 * nst, *zz = dk + 3 * nst;
    double *scale = NULL;
    if (bscale > 0.0) {
        scale = (double *)malloc(sizeof(double) * (size_t)n);
        unsigned char *vac = (unsigned char *)malloc((size_t)n);
        for (int64_t c = 0; c < n; ++c) {
            const double *v = m + 3 * c;
            vac[c] = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < 0.25;
        }
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                const int64_t c = j * nx + i;
                scale[c] = 1.0;
                if (vac[c]) continue;
                int near = 0;
                for (int64_t dj = -bmargin; dj <= bmargin && !near; ++dj)
                    for (int64_t di = -bmargin; di <= bmargin; ++di) {
                        int64_t ni = i + di, nj = j + dj;
                        if (ni < 0 || ni >= nx || nj < 0 || nj >= ny) continue;
                        if (vac[nj * nx + ni]) { near = 1; break; }
                    }
                if (!near) continue;
                unsigned n_mat = 0, n_tot = 0;
                for (int64_t dy = -r; dy <= r; ++dy) {
                    int64_t sj = j + dy;
                    if (sj < 0 || sj >= ny) continue;
                    for (int64_t dx = -r; dx <= r; ++dx) {
                        int64_t si = i + dx;
                        if (si < 0 || si >= nx) continue;
                        ++n_tot;
                        if (!vac[sj * nx + si]) ++n_mat;
                    }
                }
                double f = (double)n_mat / (double)(n_tot > 1 ? n_tot : 1);
                scale[c] = pow(f, bscale);
            }
        free(vac);
    }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            double bx = 0.0, by = 0.0, bz = 0.0;
            for (int64_t dy = -r; dy <= r; ++dy) {
                const int64_t sj = j - dy;
                if (sj < 0 || sj >= ny) continue;
                for (int64_t dx = -r; dx <= r; ++dx) {
                    const int64_t si = i - dx;
                    if (si < 0 || si >= nx) continue;
                    const int64_t k = (dy + r) * st + (dx + r);
                    const double *src = m + 3 * (sj * nx + si);
                    const double mx = ms * src[0], my = ms * src[1], mz = ms * src[2];
                    bx += xx[k] * mx + xy[k] * my;
                    by += xy[k] * mx + yy[k] * my;
                    bz += zz[k] * mz;
                }
            }
            if (scale) {
                const double s = scale[j * nx + i];
                bx *= s; by *= s; bz *= s;
            }
            double *o = b + 3 * (j * nx + i);
            o[0] += bx;
            o[1] += by;
            o[2] += bz;
        }
    free(scale);
}

/* DemagPoissonMGHybrid::add_field (:2468-2530), hybrid branch (dk may be
 * NULL when the effective radius is 0: pure MG with the screened rhs). */
void orc_hyb_add_field(orc_mg *mg, const orc_hyb_cfg *h, int64_t r, const double *dk, const double *m,
                       double *b, double ms, const orc_grid *g) {
    orc_mg_build_rhs(mg, m, ms);
    if (h->sigma_cells > 0.0) orc_mg_screen_rhs(mg, h->sigma_cells);
    orc_mg_solve(mg);          /* clears phi itself without warm start */
    orc_mg_add_b(mg, m, b);
    if (dk && r > 0) orc_hyb_add_correction(g, r, dk, m, b, ms, h->bscale, h->bmargin);
}

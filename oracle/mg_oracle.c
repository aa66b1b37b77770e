/*
 * mg_oracle.c -- CPU FP64 restatement of llg-sim's multigrid Poisson demag.
 *
 * TEST INFRASTRUCTURE ONLY (see mg_oracle.h).  Each function follows the
 * reference function cited beside it, in the same arithmetic order; rayon
 * loops become OpenMP loops with the same per-element arithmetic (reductions
 * are max-reductions, so results do not depend on the thread count).
 */
#include "mg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX3(i, j, k, nx, ny) ((((int64_t)(k)) * (ny) + (j)) * (nx) + (i))

static const double PI_ = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* configuration                                                            */
/* ------------------------------------------------------------------------ */

void orc_mg_cfg_default(orc_mg_cfg *c) {
    memset(c, 0, sizeof(*c));
    c->pad_xy = 6;
    c->n_vac_z = 16;
    c->v_cycles = 4;
    c->v_cycles_max = 64;
    c->has_tol_abs = 0;
    c->tol_abs = 0.0;
    c->has_tol_rel = 1;
    c->tol_rel = 1e-6;
    c->pre_smooth = 2;
    c->post_smooth = 2;
    c->smoother = 0;
    c->omega = 2.0 / 3.0;
    c->sor_omega = 1.0;
    c->warm_start = 1;
    c->bc = 2;
    c->tree_theta = 0.6;
    c->tree_leaf = 64;
    c->tree_max_depth = 20;
    c->stencil = 2;     /* iso27 */
    c->prolong = 1;     /* trilinear */
    c->coarse_op = 1;   /* Galerkin */
    c->iso27_alpha = 0.0;
}

/* rb_allowed_for_op / sanitize_cfg_for_op (demag_poisson_mg.rs:1400-1417) */
void orc_mg_sanitize(orc_mg_cfg *c) {
    int rb_ok = c->stencil == 0 && c->coarse_op == 0 && c->prolong == 0;
    if (c->smoother == 1 && !rb_ok) c->smoother = 0;
}

/* ------------------------------------------------------------------------ */
/* stencils                                                                 */
/* ------------------------------------------------------------------------ */

static void st_push(orc_stencil *s, int a, int b, int c, double w) {
    s->offs[s->n][0] = a;
    s->offs[s->n][1] = b;
    s->offs[s->n][2] = c;
    s->coeffs[s->n] = w;
    s->n++;
}

/* Stencil3D::seven_point (demag_poisson_mg.rs:840-872) */
void orc_stencil_seven_point(double dx, double dy, double dz, orc_stencil *s) {
    double sx = 1.0 / (dx * dx), sy = 1.0 / (dy * dy), sz = 1.0 / (dz * dz);
    s->n = 0;
    s->center = -2.0 * (sx + sy + sz);
    st_push(s, 1, 0, 0, sx);
    st_push(s, -1, 0, 0, sx);
    st_push(s, 0, 1, 0, sy);
    st_push(s, 0, -1, 0, sy);
    st_push(s, 0, 0, 1, sz);
    st_push(s, 0, 0, -1, sz);
    s->diag = -s->center;
}

/* Stencil3D::iso9_plus_z (:874-922) */
void orc_stencil_iso9_plus_z(double dx, double dy, double dz, orc_stencil *s) {
    double mx = dx > dy ? dx : dy;
    if (mx < 1e-30) mx = 1e-30;
    double rel = fabs(dx - dy) / mx;
    if (rel > 1e-6) {
        orc_stencil_seven_point(dx, dy, dz, s);
        return;
    }
    double h = 0.5 * (dx + dy);
    double inv_h2 = 1.0 / (h * h);
    double cz = 1.0 / (dz * dz);
    double c_axis = (2.0 / 3.0) * inv_h2;
    double c_diag = (1.0 / 6.0) * inv_h2;
    s->n = 0;
    s->center = -(10.0 / 3.0) * inv_h2 - 2.0 * cz;
    st_push(s, 1, 0, 0, c_axis);
    st_push(s, -1, 0, 0, c_axis);
    st_push(s, 0, 1, 0, c_axis);
    st_push(s, 0, -1, 0, c_axis);
    st_push(s, 1, 1, 0, c_diag);
    st_push(s, 1, -1, 0, c_diag);
    st_push(s, -1, 1, 0, c_diag);
    st_push(s, -1, -1, 0, c_diag);
    st_push(s, 0, 0, 1, cz);
    st_push(s, 0, 0, -1, cz);
    s->diag = -s->center;
}

/* Stencil3D::iso27 (:924-1058).  alpha <= 0 or non-finite: default alpha. */
void orc_stencil_iso27(double dx, double dy, double dz, double alpha_in, orc_stencil *s) {
    double sx1 = 0.0, sy1 = 0.0, sz1 = 0.0;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                int nn = abs(di) + abs(dj) + abs(dk);
                if (nn == 1) continue;
                double ddx = (double)di * dx, ddy = (double)dj * dy, ddz = (double)dk * dz;
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= 0.0) continue;
                double w1 = 1.0 / d2;
                sx1 += w1 * ddx * ddx;
                sy1 += w1 * ddy * ddy;
                sz1 += w1 * ddz * ddz;
            }
    int have_env = isfinite(alpha_in) && alpha_in > 0.0;
    double alpha_max = INFINITY;
    if (sx1 > 0.0) alpha_max = fmin(alpha_max, 2.0 / sx1);
    if (sy1 > 0.0) alpha_max = fmin(alpha_max, 2.0 / sy1);
    if (sz1 > 0.0) alpha_max = fmin(alpha_max, 2.0 / sz1);
    if (isfinite(alpha_max)) alpha_max *= 0.999;
    double alpha_default = isfinite(alpha_max) ? fmax(0.99 * alpha_max, 0.0) : 0.0;
    double alpha = have_env ? alpha_in : alpha_default;
    if (isfinite(alpha_max) && alpha > alpha_max) alpha = fmax(alpha_max, 0.0);

    double w_fx = (2.0 - alpha * sx1) / (2.0 * dx * dx);
    double w_fy = (2.0 - alpha * sy1) / (2.0 * dy * dy);
    double w_fz = (2.0 - alpha * sz1) / (2.0 * dz * dz);
    s->n = 0;
    st_push(s, 1, 0, 0, w_fx);
    st_push(s, -1, 0, 0, w_fx);
    st_push(s, 0, 1, 0, w_fy);
    st_push(s, 0, -1, 0, w_fy);
    st_push(s, 0, 0, 1, w_fz);
    st_push(s, 0, 0, -1, w_fz);
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                int nn = abs(di) + abs(dj) + abs(dk);
                if (nn <= 1) continue;
                double ddx = (double)di * dx, ddy = (double)dj * dy, ddz = (double)dk * dz;
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= 0.0) continue;
                st_push(s, di, dj, dk, alpha / d2);
            }
    double sum = 0.0;
    for (int q = 0; q < s->n; ++q) sum += s->coeffs[q];
    s->center = -sum;
    s->diag = sum;
}

static void stencil_from_kind(int kind, double dx, double dy, double dz, double alpha,
                              orc_stencil *s) {
    if (kind == 0) orc_stencil_seven_point(dx, dy, dz, s);
    else if (kind == 1) orc_stencil_iso9_plus_z(dx, dy, dz, s);
    else orc_stencil_iso27(dx, dy, dz, alpha, s);
}

/* offdiag_sum_at (mg_kernels.rs:72-94): neighbour indices clamped to the box */
static inline double offdiag_sum_at(const double *phi, int64_t nx, int64_t ny, int64_t nz,
                                    int64_t i, int64_t j, int64_t k, const orc_stencil *s) {
    double sum = 0.0;
    for (int q = 0; q < s->n; ++q) {
        int64_t ii = i + s->offs[q][0], jj = j + s->offs[q][1], kk = k + s->offs[q][2];
        ii = ii < 0 ? 0 : (ii > nx - 1 ? nx - 1 : ii);
        jj = jj < 0 ? 0 : (jj > ny - 1 ? ny - 1 : jj);
        kk = kk < 0 ? 0 : (kk > nz - 1 ? nz - 1 : kk);
        sum += s->coeffs[q] * phi[IDX3(ii, jj, kk, nx, ny)];
    }
    return sum;
}

/* stencil_apply_at (mg_kernels.rs:98-109) */
static inline double stencil_apply_at(const double *phi, int64_t nx, int64_t ny, int64_t nz,
                                      int64_t i, int64_t j, int64_t k, const orc_stencil *s) {
    return s->center * phi[IDX3(i, j, k, nx, ny)] + offdiag_sum_at(phi, nx, ny, nz, i, j, k, s);
}

/* apply_stencil_to_field (demag_poisson_mg.rs:1193-1211) */
void orc_mg_apply_stencil(const orc_stencil *s, const double *phi, double *out, int64_t nx,
                          int64_t ny, int64_t nz) {
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                out[IDX3(i, j, k, nx, ny)] = stencil_apply_at(phi, nx, ny, nz, i, j, k, s);
}

/* interp_1d_cell_centered (mg_kernels.rs:349-371) */
static inline void interp_1d(int64_t ic, int64_t ri, int64_t nc, int64_t r, int64_t *i0o,
                             int64_t *i1o, double *w0, double *w1) {
    int64_t i0 = ic < nc - 1 ? ic : nc - 1;
    if (r == 1) {
        *i0o = i0; *i1o = i0; *w0 = 1.0; *w1 = 0.0;
        return;
    }
    int64_t i1;
    if (ri == 0) i1 = i0 > 0 ? i0 - 1 : 0;
    else i1 = (i0 + 1) < (nc - 1) ? i0 + 1 : nc - 1;
    if (i1 == i0) { *i0o = i0; *i1o = i0; *w0 = 1.0; *w1 = 0.0; }
    else { *i0o = i0; *i1o = i1; *w0 = 0.75; *w1 = 0.25; }
}

/* prolongate_scalar (demag_poisson_mg.rs:1255-1330) */
static void prolongate_scalar(const double *c, int64_t ncx, int64_t ncy, int64_t ncz, double *f,
                              int64_t nfx, int64_t nfy, int64_t nfz, int rx, int ry, int rz,
                              int trilinear) {
    memset(f, 0, sizeof(double) * (size_t)(nfx * nfy * nfz));
    if (!trilinear) {
        for (int64_t kz = 0; kz < ncz; ++kz)
            for (int64_t jy = 0; jy < ncy; ++jy)
                for (int64_t ix = 0; ix < ncx; ++ix) {
                    double v = c[IDX3(ix, jy, kz, ncx, ncy)];
                    for (int fk = 0; fk < rz; ++fk)
                        for (int fj = 0; fj < ry; ++fj)
                            for (int fi = 0; fi < rx; ++fi)
                                f[IDX3(ix * rx + fi, jy * ry + fj, kz * rz + fk, nfx, nfy)] = v;
                }
        return;
    }
    for (int64_t k = 0; k < nfz; ++k) {
        int64_t k0, k1; double wk0, wk1;
        interp_1d(k / rz, k % rz, ncz, rz, &k0, &k1, &wk0, &wk1);
        for (int64_t j = 0; j < nfy; ++j) {
            int64_t j0, j1; double wj0, wj1;
            interp_1d(j / ry, j % ry, ncy, ry, &j0, &j1, &wj0, &wj1);
            for (int64_t i = 0; i < nfx; ++i) {
                int64_t i0, i1; double wi0, wi1;
                interp_1d(i / rx, i % rx, ncx, rx, &i0, &i1, &wi0, &wi1);
                double v = 0.0;
                v += wi0 * wj0 * wk0 * c[IDX3(i0, j0, k0, ncx, ncy)];
                v += wi1 * wj0 * wk0 * c[IDX3(i1, j0, k0, ncx, ncy)];
                v += wi0 * wj1 * wk0 * c[IDX3(i0, j1, k0, ncx, ncy)];
                v += wi1 * wj1 * wk0 * c[IDX3(i1, j1, k0, ncx, ncy)];
                v += wi0 * wj0 * wk1 * c[IDX3(i0, j0, k1, ncx, ncy)];
                v += wi1 * wj0 * wk1 * c[IDX3(i1, j0, k1, ncx, ncy)];
                v += wi0 * wj1 * wk1 * c[IDX3(i0, j1, k1, ncx, ncy)];
                v += wi1 * wj1 * wk1 * c[IDX3(i1, j1, k1, ncx, ncy)];
                f[IDX3(i, j, k, nfx, nfy)] = v;
            }
        }
    }
}

/* restrict_scalar_avg (demag_poisson_mg.rs:1213-1253) */
static void restrict_scalar_avg(const double *f, int64_t nfx, int64_t nfy, int64_t nfz, double *c,
                                int64_t ncx, int64_t ncy, int64_t ncz, int rx, int ry, int rz) {
    double norm = 1.0 / (double)(rx * ry * rz);
    for (int64_t kz = 0; kz < ncz; ++kz)
        for (int64_t jy = 0; jy < ncy; ++jy)
            for (int64_t ix = 0; ix < ncx; ++ix) {
                double sum = 0.0;
                for (int fk = 0; fk < rz; ++fk)
                    for (int fj = 0; fj < ry; ++fj)
                        for (int fi = 0; fi < rx; ++fi)
                            sum += f[IDX3(ix * rx + fi, jy * ry + fj, kz * rz + fk, nfx, nfy)];
                c[IDX3(ix, jy, kz, ncx, ncy)] = norm * sum;
            }
}

typedef struct {
    int a, b, c;
    double w;
} gal_entry;

static int gal_cmp(const void *x, const void *y) {
    const gal_entry *p = (const gal_entry *)x, *q = (const gal_entry *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->c != q->c) return p->c < q->c ? -1 : 1;
    return 0;
}

/* Stencil3D::galerkin_coarsen (demag_poisson_mg.rs:1087-1189): coarse stencil
 * R A P probed with unit vectors on a 9^3 coarse box; entries above 1e-14 in
 * magnitude kept, sorted by offset; an ill-conditioned result returns the
 * empty sentinel (n = 0, diag = 0). */
void orc_stencil_galerkin(const orc_stencil *fine, int rx, int ry, int rz, int trilinear,
                          orc_stencil *out) {
    const int64_t nc = 9;
    const int64_t nfx = nc * rx, nfy = nc * ry, nfz = nc * rz;
    const int64_t c0 = nc / 2;
    const int64_t id_c0 = IDX3(c0, c0, c0, nc, nc);
    double *phi_c = calloc((size_t)(nc * nc * nc), sizeof(double));
    double *phi_f = calloc((size_t)(nfx * nfy * nfz), sizeof(double));
    double *y_f = calloc((size_t)(nfx * nfy * nfz), sizeof(double));
    double *y_c = calloc((size_t)(nc * nc * nc), sizeof(double));
    gal_entry *ent = malloc(sizeof(gal_entry) * (size_t)(nc * nc * nc));
    int ne = 0;
    for (int64_t kz = 0; kz < nc; ++kz)
        for (int64_t jy = 0; jy < nc; ++jy)
            for (int64_t ix = 0; ix < nc; ++ix) {
                memset(phi_c, 0, sizeof(double) * (size_t)(nc * nc * nc));
                phi_c[IDX3(ix, jy, kz, nc, nc)] = 1.0;
                prolongate_scalar(phi_c, nc, nc, nc, phi_f, nfx, nfy, nfz, rx, ry, rz, trilinear);
                orc_mg_apply_stencil(fine, phi_f, y_f, nfx, nfy, nfz);
                restrict_scalar_avg(y_f, nfx, nfy, nfz, y_c, nc, nc, nc, rx, ry, rz);
                double coeff = y_c[id_c0];
                if (fabs(coeff) > 1e-14) {
                    ent[ne].a = (int)(ix - c0);
                    ent[ne].b = (int)(jy - c0);
                    ent[ne].c = (int)(kz - c0);
                    ent[ne].w = coeff;
                    ne++;
                }
            }
    qsort(ent, (size_t)ne, sizeof(gal_entry), gal_cmp);
    out->n = 0;
    out->center = 0.0;
    for (int q = 0; q < ne; ++q) {
        if (ent[q].a == 0 && ent[q].b == 0 && ent[q].c == 0) out->center = ent[q].w;
        else st_push(out, ent[q].a, ent[q].b, ent[q].c, ent[q].w);
    }
    out->diag = -out->center;
    double abs_sum = 0.0;
    int neg = 0;
    for (int q = 0; q < out->n; ++q) {
        abs_sum += fabs(out->coeffs[q]);
        if (out->coeffs[q] < 0.0) neg++;
    }
    double rho = out->diag > 0.0 ? abs_sum / out->diag : INFINITY;
    if (out->diag <= 0.0 || rho > 2.0 || neg > out->n / 2) {
        out->n = 0;
        out->center = 0.0;
        out->diag = 0.0;
    }
    free(phi_c); free(phi_f); free(y_f); free(y_c); free(ent);
}

/* ------------------------------------------------------------------------ */
/* kernels (mg_kernels.rs)                                                  */
/* ------------------------------------------------------------------------ */

/* stamp_dirichlet_bc (:35-61) */
void orc_mg_stamp_bc(double *arr, const double *bc, int64_t nx, int64_t ny, int64_t nz) {
    for (int64_t k = 0; k < nz; ++k) {
        for (int64_t j = 0; j < ny; ++j) {
            int64_t a = IDX3(0, j, k, nx, ny), b = IDX3(nx - 1, j, k, nx, ny);
            arr[a] = bc[a];
            arr[b] = bc[b];
        }
        for (int64_t i = 0; i < nx; ++i) {
            int64_t a = IDX3(i, 0, k, nx, ny), b = IDX3(i, ny - 1, k, nx, ny);
            arr[a] = bc[a];
            arr[b] = bc[b];
        }
    }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t a = IDX3(i, j, 0, nx, ny), b = IDX3(i, j, nz - 1, nx, ny);
            arr[a] = bc[a];
            arr[b] = bc[b];
        }
}

/* smooth_weighted_jacobi (:121-162) */
void orc_mg_smooth_wj(double *phi, double *tmp, const double *rhs, const double *bc, int64_t nx,
                      int64_t ny, int64_t nz, const orc_stencil *s, int64_t iters, double omega) {
    const int64_t n = nx * ny * nz;
    for (int64_t it = 0; it < iters; ++it) {
        memcpy(tmp, phi, sizeof(double) * (size_t)n);
#pragma omp parallel for schedule(static)
        for (int64_t row = 0; row < ny * nz; ++row) {
            int64_t k = row / ny, j = row % ny;
            if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) continue;
            for (int64_t i = 1; i < nx - 1; ++i) {
                int64_t id = row * nx + i;
                double off = offdiag_sum_at(tmp, nx, ny, nz, i, j, k, s);
                double gs = fabs(s->diag) > 1e-30 ? (off - rhs[id]) / s->diag : tmp[id];
                phi[id] = (1.0 - omega) * tmp[id] + omega * gs;
            }
        }
        orc_mg_stamp_bc(phi, bc, nx, ny, nz);
    }
}

/* smooth_rb_sor (:172-240) */
void orc_mg_smooth_rb(double *phi, double *tmp, const double *rhs, const double *bc, int64_t nx,
                      int64_t ny, int64_t nz, double inv_dx2, double inv_dy2, double inv_dz2,
                      int64_t iters, double omega) {
    const double sx = inv_dx2, sy = inv_dy2, sz = inv_dz2;
    const double denom = 2.0 * (sx + sy + sz);
    const int64_t plane = nx * ny;
    for (int64_t it = 0; it < iters; ++it) {
        for (int64_t color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
            for (int64_t row = 0; row < ny * nz; ++row) {
                int64_t k = row / ny, j = row % ny;
                if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) continue;
                for (int64_t i = 1; i < nx - 1; ++i) {
                    if (((i + j + k) & 1) != color) continue;
                    int64_t id = row * nx + i;
                    double off = sx * (phi[id - 1] + phi[id + 1]) + sy * (phi[id - nx] + phi[id + nx]) +
                                 sz * (phi[id - plane] + phi[id + plane]);
                    double pn = (off - rhs[id]) / denom;
                    double po = phi[id];
                    tmp[id] = po + omega * (pn - po);
                }
            }
#pragma omp parallel for schedule(static)
            for (int64_t row = 0; row < ny * nz; ++row) {
                int64_t k = row / ny, j = row % ny;
                if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) continue;
                for (int64_t i = 1; i < nx - 1; ++i) {
                    if (((i + j + k) & 1) != color) continue;
                    phi[row * nx + i] = tmp[row * nx + i];
                }
            }
        }
        orc_mg_stamp_bc(phi, bc, nx, ny, nz);
    }
}

/* compute_residual (:249-285): r = rhs - A phi on the interior, 0 on the
 * boundary; returns max |r| */
double orc_mg_residual(const double *phi, const double *rhs, double *res, int64_t nx, int64_t ny,
                       int64_t nz, const orc_stencil *s) {
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t row = 0; row < ny * nz; ++row) {
        int64_t k = row / ny, j = row % ny;
        double *r = res + row * nx;
        if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) {
            for (int64_t i = 0; i < nx; ++i) r[i] = 0.0;
            continue;
        }
        r[0] = 0.0;
        r[nx - 1] = 0.0;
        for (int64_t i = 1; i < nx - 1; ++i) {
            double a = stencil_apply_at(phi, nx, ny, nz, i, j, k, s);
            double v = rhs[row * nx + i] - a;
            r[i] = v;
            double av = fabs(v);
            if (av > mx) mx = av;
        }
    }
    return mx;
}

/* restrict_residual (:293-337) */
void orc_mg_restrict(const double *fr, int64_t fnx, int64_t fny, int64_t fnz, double *crhs,
                     double *cphi, int64_t cnx, int64_t cny, int64_t cnz) {
    const int64_t rx = fnx / cnx, ry = fny / cny, rz = fnz / cnz;
    const int64_t n = cnx * cny * cnz;
    memset(crhs, 0, sizeof(double) * (size_t)n);
    memset(cphi, 0, sizeof(double) * (size_t)n);
    const double w = 1.0 / (double)(rx * ry * rz);
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < cny * cnz; ++row) {
        int64_t kc = row / cny, jc = row % cny;
        if (kc == 0 || kc + 1 == cnz || jc == 0 || jc + 1 == cny) continue;
        for (int64_t ic = 1; ic < cnx - 1; ++ic) {
            double sum = 0.0;
            for (int64_t dk = 0; dk < rz; ++dk)
                for (int64_t dj = 0; dj < ry; ++dj)
                    for (int64_t di = 0; di < rx; ++di)
                        sum += fr[IDX3(rx * ic + di, ry * jc + dj, rz * kc + dk, fnx, fny)];
            crhs[row * cnx + ic] = w * sum;
        }
    }
}

/* prolongate_add (:377-439) */
void orc_mg_prolong_add(const double *c, int64_t cnx, int64_t cny, int64_t cnz, double *f,
                        int64_t fnx, int64_t fny, int64_t fnz, int trilinear) {
    const int64_t rx = fnx / cnx, ry = fny / cny, rz = fnz / cnz;
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < fny * fnz; ++row) {
        int64_t k = row / fny, j = row % fny;
        if (k == 0 || k + 1 == fnz || j == 0 || j + 1 == fny) continue;
        double *pr = f + row * fnx;
        if (trilinear) {
            int64_t k0, k1, j0, j1; double wk0, wk1, wj0, wj1;
            interp_1d(k / rz, k % rz, cnz, rz, &k0, &k1, &wk0, &wk1);
            interp_1d(j / ry, j % ry, cny, ry, &j0, &j1, &wj0, &wj1);
            for (int64_t i = 1; i < fnx - 1; ++i) {
                int64_t i0, i1; double wi0, wi1;
                interp_1d(i / rx, i % rx, cnx, rx, &i0, &i1, &wi0, &wi1);
                double v = wi0 * wj0 * wk0 * c[IDX3(i0, j0, k0, cnx, cny)]
                         + wi1 * wj0 * wk0 * c[IDX3(i1, j0, k0, cnx, cny)]
                         + wi0 * wj1 * wk0 * c[IDX3(i0, j1, k0, cnx, cny)]
                         + wi1 * wj1 * wk0 * c[IDX3(i1, j1, k0, cnx, cny)]
                         + wi0 * wj0 * wk1 * c[IDX3(i0, j0, k1, cnx, cny)]
                         + wi1 * wj0 * wk1 * c[IDX3(i1, j0, k1, cnx, cny)]
                         + wi0 * wj1 * wk1 * c[IDX3(i0, j1, k1, cnx, cny)]
                         + wi1 * wj1 * wk1 * c[IDX3(i1, j1, k1, cnx, cny)];
                pr[i] += v;
            }
        } else {
            int64_t kc = k / rz, jc = j / ry;
            for (int64_t i = 1; i < fnx - 1; ++i) {
                int64_t ic = i / rx;
                if (ic == 0 || ic + 1 >= cnx || jc == 0 || jc + 1 >= cny || kc == 0 || kc + 1 >= cnz)
                    continue;
                pr[i] += c[IDX3(ic, jc, kc, cnx, cny)];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* treecode and dipole boundary values (mg_treecode.rs)                    */
/* ------------------------------------------------------------------------ */

/* build_charges_from_rhs (:273-300): q = rhs dV at box-centred cell centres,
 * cells with |rhs| < 1e-40 skipped; returns the count (pos_q may be NULL). */
int64_t orc_mg_build_charges(const double *rhs, int64_t nx, int64_t ny, int64_t nz, double dx,
                             double dy, double dz, double *pos_q) {
    const double dvol = dx * dy * dz;
    const double cx = (double)nx * 0.5, cy = (double)ny * 0.5, cz = (double)nz * 0.5;
    int64_t n = 0;
    for (int64_t k = 0; k < nz; ++k) {
        double z = ((double)k + 0.5 - cz) * dz;
        for (int64_t j = 0; j < ny; ++j) {
            double y = ((double)j + 0.5 - cy) * dy;
            for (int64_t i = 0; i < nx; ++i) {
                double rho = rhs[IDX3(i, j, k, nx, ny)];
                if (fabs(rho) < 1e-40) continue;
                if (pos_q) {
                    pos_q[4 * n + 0] = ((double)i + 0.5 - cx) * dx;
                    pos_q[4 * n + 1] = y;
                    pos_q[4 * n + 2] = z;
                    pos_q[4 * n + 3] = rho * dvol;
                }
                n++;
            }
        }
    }
    return n;
}

typedef struct {
    double center[3], half, q, p[3];
    int64_t child[8];
    int64_t *idx;
    int64_t nidx, cap;
} bh_node;

typedef struct {
    const double *pq;   /* 4 per charge */
    bh_node *nodes;
    int64_t nn, capn;
    double theta;
    int64_t leaf, max_depth;
} bh_tree;

static int64_t bh_new_node(bh_tree *t, const double c[3], double half) {
    if (t->nn == t->capn) {
        t->capn = t->capn ? 2 * t->capn : 1024;
        t->nodes = realloc(t->nodes, sizeof(bh_node) * (size_t)t->capn);
    }
    bh_node *n = &t->nodes[t->nn];
    memset(n, 0, sizeof(*n));
    n->center[0] = c[0]; n->center[1] = c[1]; n->center[2] = c[2];
    n->half = half;
    for (int o = 0; o < 8; ++o) n->child[o] = -1;
    return t->nn++;
}

static int bh_is_leaf(const bh_node *n) {
    for (int o = 0; o < 8; ++o)
        if (n->child[o] >= 0) return 0;
    return 1;
}

static void bh_push(bh_node *n, int64_t ci) {
    if (n->nidx == n->cap) {
        n->cap = n->cap ? 2 * n->cap : 16;
        n->idx = realloc(n->idx, sizeof(int64_t) * (size_t)n->cap);
    }
    n->idx[n->nidx++] = ci;
}

/* BarnesHutTree::octant (:104-111) */
static int bh_octant(const double c[3], const double *pos) {
    int o = 0;
    if (pos[0] >= c[0]) o |= 1;
    if (pos[1] >= c[1]) o |= 2;
    if (pos[2] >= c[2]) o |= 4;
    return o;
}

/* ensure_child (:113-130) */
static int64_t bh_ensure_child(bh_tree *t, int64_t ni, int oct) {
    if (t->nodes[ni].child[oct] >= 0) return t->nodes[ni].child[oct];
    double pc[3] = {t->nodes[ni].center[0], t->nodes[ni].center[1], t->nodes[ni].center[2]};
    double ch = t->nodes[ni].half * 0.5;
    double c[3] = {pc[0] + ((oct & 1) ? ch : -ch), pc[1] + ((oct & 2) ? ch : -ch),
                   pc[2] + ((oct & 4) ? ch : -ch)};
    int64_t ci = bh_new_node(t, c, ch);   /* may move t->nodes */
    t->nodes[ni].child[oct] = ci;
    return ci;
}

/* insert (:132-156) */
static void bh_insert(bh_tree *t, int64_t ni, int64_t ci, int64_t depth) {
    if (depth >= t->max_depth) {
        bh_push(&t->nodes[ni], ci);
        return;
    }
    if (bh_is_leaf(&t->nodes[ni])) {
        bh_push(&t->nodes[ni], ci);
        if (t->nodes[ni].nidx > t->leaf) {
            int64_t *ids = t->nodes[ni].idx;
            int64_t cnt = t->nodes[ni].nidx;
            t->nodes[ni].idx = NULL;
            t->nodes[ni].nidx = 0;
            t->nodes[ni].cap = 0;
            for (int64_t q = 0; q < cnt; ++q) {
                int oct = bh_octant(t->nodes[ni].center, t->pq + 4 * ids[q]);
                int64_t ch = bh_ensure_child(t, ni, oct);
                bh_insert(t, ch, ids[q], depth + 1);
            }
            free(ids);
        }
    } else {
        int oct = bh_octant(t->nodes[ni].center, t->pq + 4 * ci);
        int64_t ch = bh_ensure_child(t, ni, oct);
        bh_insert(t, ch, ci, depth + 1);
    }
}

/* compute_moments_rec (:158-190): q and dipole p about the node centre */
static void bh_moments(bh_tree *t, int64_t ni, double *q_out, double p_out[3]) {
    bh_node *n = &t->nodes[ni];
    double c[3] = {n->center[0], n->center[1], n->center[2]};
    double q = 0.0, p[3] = {0.0, 0.0, 0.0};
    if (bh_is_leaf(n)) {
        for (int64_t a = 0; a < n->nidx; ++a) {
            const double *ch = t->pq + 4 * n->idx[a];
            q += ch[3];
            p[0] += ch[3] * (ch[0] - c[0]);
            p[1] += ch[3] * (ch[1] - c[1]);
            p[2] += ch[3] * (ch[2] - c[2]);
        }
    } else {
        for (int o = 0; o < 8; ++o) {
            int64_t ci = t->nodes[ni].child[o];
            if (ci < 0) continue;
            double cq, cp[3];
            bh_moments(t, ci, &cq, cp);
            const double *cc = t->nodes[ci].center;
            q += cq;
            p[0] += cp[0] + cq * (cc[0] - c[0]);
            p[1] += cp[1] + cq * (cc[1] - c[1]);
            p[2] += cp[2] + cq * (cc[2] - c[2]);
        }
    }
    t->nodes[ni].q = q;
    t->nodes[ni].p[0] = p[0];
    t->nodes[ni].p[1] = p[1];
    t->nodes[ni].p[2] = p[2];
    *q_out = q;
    p_out[0] = p[0]; p_out[1] = p[1]; p_out[2] = p[2];
}

/* eval_node (:197-259) */
static double bh_eval(const bh_tree *t, int64_t ni, const double tg[3]) {
    const bh_node *n = &t->nodes[ni];
    int leaf = bh_is_leaf(n);
    if (n->q == 0.0 && n->nidx == 0 && leaf) return 0.0;
    double rx = tg[0] - n->center[0], ry = tg[1] - n->center[1], rz = tg[2] - n->center[2];
    double ddx = fmax(fabs(rx) - n->half, 0.0), ddy = fmax(fabs(ry) - n->half, 0.0),
           ddz = fmax(fabs(rz) - n->half, 0.0);
    double d = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    double size = n->half * 2.0;
    int accept = leaf ? 1 : (d > 0.0 ? (size / d < t->theta) : 0);
    if (accept) {
        if (leaf) {
            double sum = 0.0;
            for (int64_t a = 0; a < n->nidx; ++a) {
                const double *c = t->pq + 4 * n->idx[a];
                double ex = tg[0] - c[0], ey = tg[1] - c[1], ez = tg[2] - c[2];
                double rr2 = ex * ex + ey * ey + ez * ez;
                if (rr2 > 0.0) sum += c[3] / sqrt(rr2);
            }
            return sum;
        }
        double inv_r = 1.0 / r;
        double inv_r3 = inv_r * inv_r * inv_r;
        double pr = n->p[0] * rx + n->p[1] * ry + n->p[2] * rz;
        return n->q * inv_r + pr * inv_r3;
    }
    double sum = 0.0;
    for (int o = 0; o < 8; ++o)
        if (n->child[o] >= 0) sum += bh_eval(t, n->child[o], tg);
    return sum;
}

static void bh_build(bh_tree *t, const double *pq, int64_t nc, const double rc[3], double half,
                     int64_t leaf, double theta, int64_t max_depth) {
    memset(t, 0, sizeof(*t));
    t->pq = pq;
    t->theta = theta;
    t->leaf = leaf;
    t->max_depth = max_depth;
    bh_new_node(t, rc, half);
    for (int64_t a = 0; a < nc; ++a) bh_insert(t, 0, a, 0);
    double q, p[3];
    bh_moments(t, 0, &q, p);
}

static void bh_free(bh_tree *t) {
    for (int64_t a = 0; a < t->nn; ++a) free(t->nodes[a].idx);
    free(t->nodes);
}

/* BarnesHutTree::build + eval_phi on arbitrary targets (3 per target) */
void orc_mg_treecode_eval(const double *pos_q, int64_t n_charges, const double root_center[3],
                          double root_half, int64_t leaf, double theta, int64_t max_depth,
                          const double *targets, int64_t n_targets, double *phi_out) {
    bh_tree t;
    bh_build(&t, pos_q, n_charges, root_center, root_half, leaf, theta, max_depth);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t a = 0; a < n_targets; ++a)
        phi_out[a] = -bh_eval(&t, 0, targets + 3 * a) / (4.0 * PI_);
    bh_free(&t);
}

/* evaluate_treecode_bc (:308-366) with the charges of build_charges_from_rhs */
void orc_mg_treecode_bc(double *bc, const double *rhs, int64_t nx, int64_t ny, int64_t nz,
                        double dx, double dy, double dz, int64_t leaf, double theta,
                        int64_t max_depth) {
    memset(bc, 0, sizeof(double) * (size_t)(nx * ny * nz));
    int64_t nc = orc_mg_build_charges(rhs, nx, ny, nz, dx, dy, dz, NULL);
    if (nc == 0) return;
    double *pq = malloc(sizeof(double) * 4 * (size_t)nc);
    orc_mg_build_charges(rhs, nx, ny, nz, dx, dy, dz, pq);
    double lx = (double)nx * dx, ly = (double)ny * dy, lz = (double)nz * dz;
    double root_half = 0.5 * fmax(fmax(lx, ly), lz) + 1e-12;
    double cx = (double)nx * 0.5, cy = (double)ny * 0.5, cz = (double)nz * 0.5;
    const double rc[3] = {0.0, 0.0, 0.0};
    bh_tree t;
    bh_build(&t, pq, nc, rc, root_half, leaf, theta, max_depth);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t row = 0; row < ny * nz; ++row) {
        int64_t k = row / ny, j = row % ny;
        double z = ((double)k + 0.5 - cz) * dz;
        double y = ((double)j + 0.5 - cy) * dy;
        double *br = bc + row * nx;
        if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) {
            for (int64_t i = 0; i < nx; ++i) {
                double tg[3] = {((double)i + 0.5 - cx) * dx, y, z};
                br[i] = -bh_eval(&t, 0, tg) / (4.0 * PI_);
            }
        } else {
            double t0[3] = {(0.5 - cx) * dx, y, z};
            br[0] = -bh_eval(&t, 0, t0) / (4.0 * PI_);
            double t1[3] = {((double)nx - 0.5 - cx) * dx, y, z};
            br[nx - 1] = -bh_eval(&t, 0, t1) / (4.0 * PI_);
        }
    }
    bh_free(&t);
    free(pq);
}

/* evaluate_dipole_bc (:371-448) */
void orc_mg_dipole_bc(double *bc, const double *rhs, int64_t nx, int64_t ny, int64_t nz,
                      double dx, double dy, double dz) {
    memset(bc, 0, sizeof(double) * (size_t)(nx * ny * nz));
    const double dvol = dx * dy * dz;
    const double cx = (double)nx * 0.5, cy = (double)ny * 0.5, cz = (double)nz * 0.5;
    double q = 0.0, pxm = 0.0, pym = 0.0, pzm = 0.0;
    for (int64_t k = 0; k < nz; ++k) {
        double z = ((double)k + 0.5 - cz) * dz;
        for (int64_t j = 0; j < ny; ++j) {
            double y = ((double)j + 0.5 - cy) * dy;
            for (int64_t i = 0; i < nx; ++i) {
                double x = ((double)i + 0.5 - cx) * dx;
                double w = rhs[IDX3(i, j, k, nx, ny)] * dvol;
                q += w;
                pxm += w * x;
                pym += w * y;
                pzm += w * z;
            }
        }
    }
    const double inv4pi = 1.0 / (4.0 * PI_);
#define DIP_EVAL(X, Y, Z, OUT)                                                       \
    do {                                                                             \
        double x_ = (X), y_ = (Y), z_ = (Z);                                         \
        double r2_ = x_ * x_ + y_ * y_ + z_ * z_;                                    \
        if (r2_ <= 0.0) { (OUT) = 0.0; break; }                                      \
        double r_ = sqrt(r2_);                                                       \
        double pr_ = pxm * x_ + pym * y_ + pzm * z_;                                 \
        (OUT) = -inv4pi * (q / r_ + pr_ / (r_ * r_ * r_));                           \
    } while (0)
    for (int64_t k = 0; k < nz; ++k) {
        double z = ((double)k + 0.5 - cz) * dz;
        for (int64_t j = 0; j < ny; ++j) {
            double y = ((double)j + 0.5 - cy) * dy;
            DIP_EVAL((0.5 - cx) * dx, y, z, bc[IDX3(0, j, k, nx, ny)]);
            DIP_EVAL(((double)nx - 0.5 - cx) * dx, y, z, bc[IDX3(nx - 1, j, k, nx, ny)]);
        }
        for (int64_t i = 0; i < nx; ++i) {
            double x = ((double)i + 0.5 - cx) * dx;
            DIP_EVAL(x, (0.5 - cy) * dy, z, bc[IDX3(i, 0, k, nx, ny)]);
            DIP_EVAL(x, ((double)ny - 0.5 - cy) * dy, z, bc[IDX3(i, ny - 1, k, nx, ny)]);
        }
    }
    for (int64_t j = 0; j < ny; ++j) {
        double y = ((double)j + 0.5 - cy) * dy;
        for (int64_t i = 0; i < nx; ++i) {
            double x = ((double)i + 0.5 - cx) * dx;
            DIP_EVAL(x, y, (0.5 - cz) * dz, bc[IDX3(i, j, 0, nx, ny)]);
            DIP_EVAL(x, y, ((double)nz - 0.5 - cz) * dz, bc[IDX3(i, j, nz - 1, nx, ny)]);
        }
    }
#undef DIP_EVAL
}

/* ------------------------------------------------------------------------ */
/* solver                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t nx, ny, nz;
    double dx, dy, dz, inv_dx2, inv_dy2, inv_dz2;
    orc_stencil st;
    double *phi, *rhs, *res, *tmp, *bc;
} mg_level;

struct orc_mg {
    orc_grid g;
    orc_mg_cfg cfg;
    int64_t px, py, pz, offx, offy, offz;
    int nlev;
    mg_level lev[ORC_MG_MAX_LEVELS];
    orc_mg_stats stats;
};

/* MGLevel::new (demag_poisson_mg.rs:1354-1374) */
static void level_init(mg_level *L, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                       double dz) {
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->dx = dx; L->dy = dy; L->dz = dz;
    L->inv_dx2 = 1.0 / (dx * dx);
    L->inv_dy2 = 1.0 / (dy * dy);
    L->inv_dz2 = 1.0 / (dz * dz);
    orc_stencil_seven_point(dx, dy, dz, &L->st);
    size_t n = (size_t)(nx * ny * nz);
    L->phi = calloc(n, sizeof(double));
    L->rhs = calloc(n, sizeof(double));
    L->res = calloc(n, sizeof(double));
    L->tmp = calloc(n, sizeof(double));
    L->bc = calloc(n, sizeof(double));
}

/* new_with_operator (:1425-1577) */
orc_mg *orc_mg_new(const orc_grid *g, const orc_mg_cfg *cfg_in) {
    orc_mg *s = calloc(1, sizeof(orc_mg));
    s->g = *g;
    s->cfg = *cfg_in;
    orc_mg_sanitize(&s->cfg);
    const orc_mg_cfg *cfg = &s->cfg;
    int64_t nx = g->nx > 1 ? g->nx : 1, ny = g->ny > 1 ? g->ny : 1;
    int64_t pad = cfg->pad_xy > 1 ? cfg->pad_xy : 1;
    int64_t px = nx + 2 * pad, py = ny + 2 * pad;
    if (px % 2 == 1) px++;
    if (py % 2 == 1) py++;
    int64_t n_vac = cfg->n_vac_z > 1 ? cfg->n_vac_z : 1;
    int64_t pz = 1 + 2 * n_vac;
    if (pz % 2 == 1) pz++;
    if (pz % 4 != 0) pz = ((pz + 3) / 4) * 4;
    s->px = px; s->py = py; s->pz = pz;
    s->offx = px > nx ? (px - nx) / 2 : 0;
    s->offy = py > ny ? (py - ny) / 2 : 0;
    s->offz = n_vac;

    int64_t lx = px, ly = py, lz = pz;
    double dx = g->dx, dy = g->dy, dz = g->dz;
    level_init(&s->lev[0], lx, ly, lz, dx, dy, dz);
    s->nlev = 1;
    for (;;) {
        int can_xy = lx >= 8 && ly >= 8 && lx % 2 == 0 && ly % 2 == 0;
        int can_z = lz >= 8 && lz % 2 == 0;
        if (!can_xy && !can_z) break;
        if (can_xy) { lx /= 2; ly /= 2; dx *= 2.0; dy *= 2.0; }
        if (can_z) { lz /= 2; dz *= 2.0; }
        level_init(&s->lev[s->nlev], lx, ly, lz, dx, dy, dz);
        s->nlev++;
        if (s->nlev > 32) break;
    }
    stencil_from_kind(cfg->stencil, s->lev[0].dx, s->lev[0].dy, s->lev[0].dz, cfg->iso27_alpha,
                      &s->lev[0].st);
    for (int l = 1; l < s->nlev; ++l) {
        mg_level *F = &s->lev[l - 1], *Cl = &s->lev[l];
        int rx = (int)(F->nx / Cl->nx), ry = (int)(F->ny / Cl->ny), rz = (int)(F->nz / Cl->nz);
        if (cfg->coarse_op == 0) {
            stencil_from_kind(cfg->stencil, Cl->dx, Cl->dy, Cl->dz, cfg->iso27_alpha, &Cl->st);
        } else {
            orc_stencil_galerkin(&F->st, rx, ry, rz, cfg->prolong == 1, &Cl->st);
            if (Cl->st.diag <= 0.0 || Cl->st.n == 0)
                stencil_from_kind(cfg->stencil, Cl->dx, Cl->dy, Cl->dz, cfg->iso27_alpha, &Cl->st);
        }
    }
    s->stats.last_residual = -1.0;
    return s;
}

void orc_mg_free(orc_mg *s) {
    if (!s) return;
    for (int l = 0; l < s->nlev; ++l) {
        free(s->lev[l].phi); free(s->lev[l].rhs); free(s->lev[l].res);
        free(s->lev[l].tmp); free(s->lev[l].bc);
    }
    free(s);
}

/* same_structure (:1585-1593) */
int orc_mg_same_structure(const orc_mg *s, const orc_grid *g, const orc_mg_cfg *c) {
    return s->g.nx == g->nx && s->g.ny == g->ny && s->g.dx == g->dx && s->g.dy == g->dy &&
           s->g.dz == g->dz && s->cfg.pad_xy == c->pad_xy && s->cfg.n_vac_z == c->n_vac_z &&
           s->cfg.stencil == c->stencil && s->cfg.prolong == c->prolong &&
           s->cfg.coarse_op == c->coarse_op &&
           (s->cfg.iso27_alpha == c->iso27_alpha ||
            (!(s->cfg.iso27_alpha > 0.0) && !(c->iso27_alpha > 0.0)));
}

/* apply_cfg (:1579-1583): runtime knobs; the operator stays as built */
void orc_mg_apply_cfg(orc_mg *s, const orc_mg_cfg *c) {
    orc_mg_cfg k = *c;
    k.stencil = s->cfg.stencil;
    k.prolong = s->cfg.prolong;
    k.coarse_op = s->cfg.coarse_op;
    k.iso27_alpha = s->cfg.iso27_alpha;
    orc_mg_sanitize(&k);
    s->cfg = k;
}

void orc_mg_dims(const orc_mg *s, int64_t d[6]) {
    d[0] = s->px; d[1] = s->py; d[2] = s->pz;
    d[3] = s->offx; d[4] = s->offy; d[5] = s->offz;
}

int orc_mg_levels(const orc_mg *s, int64_t *dims, double *sp) {
    for (int l = 0; l < s->nlev; ++l) {
        if (dims) { dims[3 * l] = s->lev[l].nx; dims[3 * l + 1] = s->lev[l].ny; dims[3 * l + 2] = s->lev[l].nz; }
        if (sp) { sp[3 * l] = s->lev[l].dx; sp[3 * l + 1] = s->lev[l].dy; sp[3 * l + 2] = s->lev[l].dz; }
    }
    return s->nlev;
}

void orc_mg_level_stencil(const orc_mg *s, int level, orc_stencil *out) { *out = s->lev[level].st; }

double *orc_mg_finest(orc_mg *s, int which) {
    mg_level *L = &s->lev[0];
    return which == 0 ? L->phi : which == 1 ? L->rhs : which == 2 ? L->res : L->bc;
}

void orc_mg_stats_get(const orc_mg *s, orc_mg_stats *out) { *out = s->stats; }

/* m_at (demag_poisson_mg.rs:1620-1662) */
static inline int m_at(const orc_mg *s, int64_t pi, int64_t pj, int64_t pk, const double *m,
                       double ms, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (pi < 0 || pj < 0 || pk < 0) return 0;
    if (pi >= s->px || pj >= s->py || pk >= s->pz) return 0;
    if (pk != s->offz) return 0;
    if (pi < s->offx || pi >= s->offx + s->g.nx || pj < s->offy || pj >= s->offy + s->g.ny) return 0;
    const double *v = m + 3 * ((pj - s->offy) * s->g.nx + (pi - s->offx));
    double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (n2 < 1e-30) return 0;
    out[0] = ms * v[0];
    out[1] = ms * v[1];
    out[2] = ms * v[2];
    return 1;
}

/* face_val (:1664-1672) */
static inline double face_val(int ia, double a, int ib, double b) {
    if (ia && ib) return 0.5 * (a + b);
    if (ia) return a;
    if (ib) return b;
    return 0.0;
}

/* build_rhs_from_m (:1595-1802): rhs = div(Ms m) with face-averaged values */
void orc_mg_build_rhs(orc_mg *s, const double *m, double ms) {
    mg_level *L = &s->lev[0];
    const int64_t nx = L->nx, ny = L->ny, nz = L->nz;
    const double dx = L->dx, dy = L->dy, dz = L->dz;
    memset(L->rhs, 0, sizeof(double) * (size_t)(nx * ny * nz));
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < ny * nz; ++row) {
        int64_t k = row / ny, j = row % ny;
        if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) continue;
        for (int64_t i = 1; i < nx - 1; ++i) {
            double c[3], xp[3], xm[3], yp[3], ym[3], zp[3], zm[3];
            int ci = m_at(s, i, j, k, m, ms, c);
            int xpi = m_at(s, i + 1, j, k, m, ms, xp);
            int xmi = m_at(s, i - 1, j, k, m, ms, xm);
            int ypi = m_at(s, i, j + 1, k, m, ms, yp);
            int ymi = m_at(s, i, j - 1, k, m, ms, ym);
            int zpi = m_at(s, i, j, k + 1, m, ms, zp);
            int zmi = m_at(s, i, j, k - 1, m, ms, zm);
            double mx_p = face_val(ci, c[0], xpi, xp[0]);
            double mx_m = face_val(xmi, xm[0], ci, c[0]);
            double my_p = face_val(ci, c[1], ypi, yp[1]);
            double my_m = face_val(ymi, ym[1], ci, c[1]);
            double mz_p = face_val(ci, c[2], zpi, zp[2]);
            double mz_m = face_val(zmi, zm[2], ci, c[2]);
            L->rhs[row * nx + i] = (mx_p - mx_m) / dx + (my_p - my_m) / dy + (mz_p - mz_m) / dz;
        }
    }
}

/* update_finest_boundary_bc (:1804-1833) */
static void update_bc(orc_mg *s) {
    mg_level *L = &s->lev[0];
    s->stats.n_charges = 0;
    if (s->cfg.bc == 0) {
        memset(L->bc, 0, sizeof(double) * (size_t)(L->nx * L->ny * L->nz));
    } else if (s->cfg.bc == 1) {
        orc_mg_dipole_bc(L->bc, L->rhs, L->nx, L->ny, L->nz, L->dx, L->dy, L->dz);
    } else {
        s->stats.n_charges =
            (int32_t)orc_mg_build_charges(L->rhs, L->nx, L->ny, L->nz, L->dx, L->dy, L->dz, NULL);
        orc_mg_treecode_bc(L->bc, L->rhs, L->nx, L->ny, L->nz, L->dx, L->dy, L->dz,
                           s->cfg.tree_leaf, s->cfg.tree_theta, s->cfg.tree_max_depth);
    }
    orc_mg_stamp_bc(L->phi, L->bc, L->nx, L->ny, L->nz);
}

static void smooth(orc_mg *s, int l, int64_t iters) {
    mg_level *L = &s->lev[l];
    if (s->cfg.smoother == 0)
        orc_mg_smooth_wj(L->phi, L->tmp, L->rhs, L->bc, L->nx, L->ny, L->nz, &L->st, iters,
                         s->cfg.omega);
    else
        orc_mg_smooth_rb(L->phi, L->tmp, L->rhs, L->bc, L->nx, L->ny, L->nz, L->inv_dx2,
                         L->inv_dy2, L->inv_dz2, iters, s->cfg.sor_omega);
}

/* v_cycle (:1877-1945) */
static void v_cycle(orc_mg *s, int l) {
    if (l == s->nlev - 1) {
        smooth(s, l, 80);
        return;
    }
    smooth(s, l, s->cfg.pre_smooth);
    mg_level *F = &s->lev[l], *Cl = &s->lev[l + 1];
    orc_mg_residual(F->phi, F->rhs, F->res, F->nx, F->ny, F->nz, &F->st);
    orc_mg_restrict(F->res, F->nx, F->ny, F->nz, Cl->rhs, Cl->phi, Cl->nx, Cl->ny, Cl->nz);
    v_cycle(s, l + 1);
    orc_mg_prolong_add(Cl->phi, Cl->nx, Cl->ny, Cl->nz, F->phi, F->nx, F->ny, F->nz,
                       s->cfg.prolong == 1);
    smooth(s, l, s->cfg.post_smooth);
}

/* solve_with_timing (:1947-2057) */
void orc_mg_solve(orc_mg *s) {
    mg_level *L = &s->lev[0];
    const int64_t n = L->nx * L->ny * L->nz;
    if (!s->cfg.warm_start) memset(L->phi, 0, sizeof(double) * (size_t)n);
    update_bc(s);
    orc_mg_stamp_bc(L->phi, L->bc, L->nx, L->ny, L->nz);

    const int use_tol = s->cfg.has_tol_abs || s->cfg.has_tol_rel;
    double rhs_max = 0.0;
    if (use_tol && s->cfg.has_tol_rel) {
        for (int64_t row = 0; row < L->ny * L->nz; ++row) {
            int64_t k = row / L->ny, j = row % L->ny;
            if (k == 0 || k + 1 == L->nz || j == 0 || j + 1 == L->ny) continue;
            for (int64_t i = 1; i < L->nx - 1; ++i) {
                double v = fabs(L->rhs[row * L->nx + i]);
                if (v > rhs_max) rhs_max = v;
            }
        }
    }
    double tol_target = 0.0;
    if (use_tol) {
        if (s->cfg.has_tol_abs) tol_target = fmax(tol_target, fmax(s->cfg.tol_abs, 0.0));
        if (s->cfg.has_tol_rel) tol_target = fmax(tol_target, fmax(fmax(s->cfg.tol_rel, 0.0) * rhs_max, 0.0));
    }
    int64_t min_cycles = use_tol ? (s->cfg.v_cycles > 1 ? s->cfg.v_cycles : 1) : 0;
    int64_t max_cycles;
    if (use_tol) {
        max_cycles = s->cfg.v_cycles_max > min_cycles ? s->cfg.v_cycles_max : min_cycles;
        if (max_cycles < 1) max_cycles = 1;
    } else {
        max_cycles = s->cfg.v_cycles > 1 ? s->cfg.v_cycles : 1;
    }
    s->stats.rhs_max = rhs_max;
    s->stats.tol_target = tol_target;
    s->stats.last_residual = -1.0;
    s->stats.diverged = 0;
    s->stats.cycles = 0;
    for (int64_t iter = 0; iter < max_cycles; ++iter) {
        v_cycle(s, 0);
        orc_mg_stamp_bc(L->phi, L->bc, L->nx, L->ny, L->nz);
        s->stats.cycles = iter + 1;
        if (iter % 2 == 1 || iter + 1 == max_cycles) {
            int bad = 0;
            for (int64_t a = 0; a < L->nx * L->ny * 2 && a < n; ++a)
                if (!isfinite(L->phi[a])) { bad = 1; break; }
            if (bad) {
                memset(L->phi, 0, sizeof(double) * (size_t)n);
                orc_mg_stamp_bc(L->phi, L->bc, L->nx, L->ny, L->nz);
                s->stats.diverged = 1;
                break;
            }
        }
        if (use_tol && iter + 1 >= min_cycles) {
            double r = orc_mg_residual(L->phi, L->rhs, L->res, L->nx, L->ny, L->nz, &L->st);
            s->stats.last_residual = r;
            if (r <= tol_target) break;
        }
    }
    orc_mg_stamp_bc(L->phi, L->bc, L->nx, L->ny, L->nz);
}

/* add_b_from_phi_on_magnet_layer -> extract_gradient_on_magnet_layer
 * (demag_poisson_mg.rs:2063-2092, mg_kernels.rs:450-507) */
void orc_mg_add_b(const orc_mg *s, const double *m, double *b) { orc_mg_add_b_opt(s, m, b); }

/* extract_gradient_on_magnet_layer with an optional magnet mask (m NULL: every
 * film cell, add_b_from_phi_on_magnet_layer_all, demag_poisson_mg.rs:2072-2074) */
void orc_mg_add_b_opt(const orc_mg *s, const double *m, double *b) {
    const mg_level *L = &s->lev[0];
    const double *phi = L->phi;
    const int64_t px = s->px, py = s->py, pz = s->pz, k = s->offz;
    const double dx = L->dx, dy = L->dy, dz = L->dz;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < s->g.ny; ++j) {
        int64_t pj = s->offy + j;
        for (int64_t i = 0; i < s->g.nx; ++i) {
            if (m) {
                const double *mv = m + 3 * (j * s->g.nx + i);
                if (!(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2] > 1e-30)) continue;
            }
            int64_t pi = s->offx + i;
            double pc = phi[IDX3(pi, pj, k, px, py)];
            double gx = 0.0, wx = 0.0;
            if (pi + 1 < px) { gx += (phi[IDX3(pi + 1, pj, k, px, py)] - pc) / dx; wx += 1.0; }
            if (pi > 0) { gx += (pc - phi[IDX3(pi - 1, pj, k, px, py)]) / dx; wx += 1.0; }
            if (wx > 0.0) gx /= wx;
            double gy = 0.0, wy = 0.0;
            if (pj + 1 < py) { gy += (phi[IDX3(pi, pj + 1, k, px, py)] - pc) / dy; wy += 1.0; }
            if (pj > 0) { gy += (pc - phi[IDX3(pi, pj - 1, k, px, py)]) / dy; wy += 1.0; }
            if (wy > 0.0) gy /= wy;
            double gz = 0.0, wz = 0.0;
            if (k + 1 < pz) { gz += (phi[IDX3(pi, pj, k + 1, px, py)] - pc) / dz; wz += 1.0; }
            if (k > 0) { gz += (pc - phi[IDX3(pi, pj, k - 1, px, py)]) / dz; wz += 1.0; }
            if (wz > 0.0) gz /= wz;
            double *bo = b + 3 * (j * s->g.nx + i);
            bo[0] += -ORC_MU0 * gx;
            bo[1] += -ORC_MU0 * gy;
            bo[2] += -ORC_MU0 * gz;
        }
    }
}

/* DemagPoissonMG::add_field (:2094-2098) */
void orc_mg_add_field(orc_mg *s, const double *m, double *b, double ms) {
    orc_mg_build_rhs(s, m, ms);
    orc_mg_solve(s);
    orc_mg_add_b(s, m, b);
}

/* gaussian_kernel_1d (mg_kernels.rs:580-599); returns r, fills w[2r+1] */
int64_t orc_gaussian_kernel_1d(double sigma, double *w, int64_t cap) {
    if (!(sigma > 0.0)) { if (cap > 0) w[0] = 1.0; return 0; }
    int64_t r = (int64_t)ceil(3.0 * sigma);
    if (r <= 0) { if (cap > 0) w[0] = 1.0; return 0; }
    if (2 * r + 1 > cap) return -1;
    double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    for (int64_t i = -r; i <= r; ++i) {
        double x = (double)i;
        w[i + r] = exp(-x * x * inv2s2);
    }
    double sum = 0.0;
    for (int64_t i = 0; i < 2 * r + 1; ++i) sum += w[i];
    if (sum > 0.0)
        for (int64_t i = 0; i < 2 * r + 1; ++i) w[i] /= sum;
    return r;
}

/* screen_rhs_gaussian_xy (mg_kernels.rs:604-680) on the finest level's rhs,
 * its tmp array as scratch (demag_poisson_mg.rs:2714-2720) */
void orc_mg_screen_rhs(orc_mg *s, double sigma) {
    if (!(sigma > 0.0)) return;
    double w[1024];
    int64_t r = orc_gaussian_kernel_1d(sigma, w, 1024);
    if (r <= 0) return;
    mg_level *L = &s->lev[0];
    const int64_t nx = L->nx, ny = L->ny, nz = L->nz;
    if (nx < 3 || ny < 3 || nz < 3) return;
    double *rhs = L->rhs, *tmp = L->tmp;
    for (int64_t row = 0; row < ny * nz; ++row) {
        int64_t k = row / ny, j = row - k * ny;
        double *t = tmp + row * nx;
        if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) {
            for (int64_t i = 0; i < nx; ++i) t[i] = 0.0;
            continue;
        }
        const double *src = rhs + row * nx;
        t[0] = 0.0;
        t[nx - 1] = 0.0;
        for (int64_t i = 1; i < nx - 1; ++i) {
            double acc = 0.0;
            for (int64_t q = -r; q <= r; ++q) {
                int64_t x = i + q;
                if (x < 0) x = 0; else if (x > nx - 1) x = nx - 1;
                acc += w[q + r] * src[x];
            }
            t[i] = acc;
        }
    }
    for (int64_t row = 0; row < ny * nz; ++row) {
        int64_t k = row / ny, j = row - k * ny;
        double *o = rhs + row * nx;
        if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny) {
            for (int64_t i = 0; i < nx; ++i) o[i] = 0.0;
            continue;
        }
        o[0] = 0.0;
        o[nx - 1] = 0.0;
        for (int64_t i = 1; i < nx - 1; ++i) {
            double acc = 0.0;
            for (int64_t q = -r; q <= r; ++q) {
                int64_t y = j + q;
                if (y < 0) y = 0; else if (y > ny - 1) y = ny - 1;
                acc += w[q + r] * tmp[(k * ny + y) * nx + i];
            }
            o[i] = acc;
        }
    }
}

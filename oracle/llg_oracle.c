/*
 * llg_oracle.c -- CPU FP64 restatement of llg-sim's uniform-grid LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline that bench.py times.  The product library never links or calls it.
 *
 * Each function names the reference function (file:line in the llg-sim
 * source tree) whose algorithm and floating-point order it restates.  Where
 * the reference parallelises with rayon this file uses OpenMP over the same
 * independent units (rows, columns, displacements); the reductions on the
 * path are max-reductions or per-element, so results do not depend on the
 * thread count.
 *
 * The FFT stands in for the third-party crate rustfft 6.4.1 (Cargo.lock:1224):
 * it is an unnormalised DFT of any length (Stockham autosort, mixed radix
 * 4/2/3/5 and a direct DFT for other prime factors) and is pinned against the
 * DFT definition and numpy.fft, not against rustfft's rounding.
 */
#include "llg_oracle.h"
#include "mg_oracle.h"
#include "hybrid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* threads                                                                   */
/* ------------------------------------------------------------------------ */
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* vec3.rs:5-28                                                              */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
/* normalize(0) = (0,0,1) (vec3.rs:21-28) */
static inline void normalize3(const double *v, double *o) {
    double n2 = dot3(v, v);
    if (n2 == 0.0) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 1.0;
        return;
    }
    double inv = 1.0 / sqrt(n2);
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

/* ------------------------------------------------------------------------ */
/* zeeman.rs:6-11 -- every cell, vacuum included                            */
/* ------------------------------------------------------------------------ */
void orc_add_zeeman(int64_t n, double *b, const double b_ext[3]) {
    for (int64_t i = 0; i < n; ++i) {
        b[3 * i + 0] += b_ext[0];
        b[3 * i + 1] += b_ext[1];
        b[3 * i + 2] += b_ext[2];
    }
}

/* ------------------------------------------------------------------------ */
/* anisotropy.rs:10-25                                                       */
/* ------------------------------------------------------------------------ */
void orc_add_anisotropy(int64_t n, const double *m, double *b, const orc_material *mat) {
    double k_u = mat->k_u, ms = mat->ms;
    if (k_u == 0.0 || ms == 0.0) return;
    double coeff = 2.0 * k_u / ms;
    const double *u = mat->easy_axis;
    for (int64_t i = 0; i < n; ++i) {
        const double *mc = m + 3 * i;
        double mdotu = mc[0] * u[0] + mc[1] * u[1] + mc[2] * u[2];
        b[3 * i + 0] += coeff * mdotu * u[0];
        b[3 * i + 1] += coeff * mdotu * u[1];
        b[3 * i + 2] += coeff * mdotu * u[2];
    }
}

/* ------------------------------------------------------------------------ */
/* exchange.rs:8-35 (ghosts), dmi.rs:18-52 (identical ghost formulas)        */
/* ------------------------------------------------------------------------ */
static inline void ghost_x(const double *m, double n_x, double eta, double dx, double *o) {
    double dmx_dx = -eta * m[2];
    double dmy_dx = 0.0;
    double dmz_dx = eta * m[0];
    o[0] = m[0] + n_x * dx * dmx_dx;
    o[1] = m[1] + n_x * dx * dmy_dx;
    o[2] = m[2] + n_x * dx * dmz_dx;
}
static inline void ghost_y(const double *m, double n_y, double eta, double dy, double *o) {
    double dmx_dy = 0.0;
    double dmy_dy = -eta * m[2];
    double dmz_dy = eta * m[1];
    o[0] = m[0] + n_y * dy * dmx_dy;
    o[1] = m[1] + n_y * dy * dmy_dy;
    o[2] = m[2] + n_y * dy * dmz_dy;
}
static inline void copy3(const double *a, double *o) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }

/* exchange.rs:45-222.  Neighbour rules:
 *  - nx == 1: neighbour = centre;
 *  - domain edge: chiral ghost if D != 0 else centre;
 *  - masked-out neighbour: chiral ghost if D != 0 else centre;
 *  - masked-out centre: no contribution.                                   */
void orc_add_exchange(const orc_grid *g, const double *m, double *b,
                      const orc_material *mat, const uint8_t *mask) {
    int64_t nx = g->nx, ny = g->ny;
    if (nx == 0 || ny == 0) return;
    double a = mat->a_ex, ms = mat->ms;
    if (a == 0.0 || ms == 0.0) return;
    double dx2 = g->dx * g->dx, dy2 = g->dy * g->dy;
    double coeff = 2.0 * a / ms;
    int has_eta = mat->has_dmi && mat->dmi != 0.0;
    double eta = has_eta ? mat->dmi / (2.0 * a) : 0.0;

    for (int64_t j = 0; j < ny; ++j) {
        for (int64_t i = 0; i < nx; ++i) {
            int64_t c = j * nx + i;
            if (mask && !mask[c]) continue;
            const double *mij = m + 3 * c;
            double mim[3], mip[3], mjm[3], mjp[3];

            if (nx == 1) copy3(mij, mim);
            else if (i == 0) { if (has_eta) ghost_x(mij, -1.0, eta, g->dx, mim); else copy3(mij, mim); }
            else {
                int64_t l = c - 1;
                if (mask && !mask[l]) { if (has_eta) ghost_x(mij, -1.0, eta, g->dx, mim); else copy3(mij, mim); }
                else copy3(m + 3 * l, mim);
            }
            if (nx == 1) copy3(mij, mip);
            else if (i == nx - 1) { if (has_eta) ghost_x(mij, 1.0, eta, g->dx, mip); else copy3(mij, mip); }
            else {
                int64_t r = c + 1;
                if (mask && !mask[r]) { if (has_eta) ghost_x(mij, 1.0, eta, g->dx, mip); else copy3(mij, mip); }
                else copy3(m + 3 * r, mip);
            }
            if (ny == 1) copy3(mij, mjm);
            else if (j == 0) { if (has_eta) ghost_y(mij, -1.0, eta, g->dy, mjm); else copy3(mij, mjm); }
            else {
                int64_t d = c - nx;
                if (mask && !mask[d]) { if (has_eta) ghost_y(mij, -1.0, eta, g->dy, mjm); else copy3(mij, mjm); }
                else copy3(m + 3 * d, mjm);
            }
            if (ny == 1) copy3(mij, mjp);
            else if (j == ny - 1) { if (has_eta) ghost_y(mij, 1.0, eta, g->dy, mjp); else copy3(mij, mjp); }
            else {
                int64_t u = c + nx;
                if (mask && !mask[u]) { if (has_eta) ghost_y(mij, 1.0, eta, g->dy, mjp); else copy3(mij, mjp); }
                else copy3(m + 3 * u, mjp);
            }
            for (int k = 0; k < 3; ++k) {
                double d2x = nx > 1 ? (mip[k] - 2.0 * mij[k] + mim[k]) / dx2 : 0.0;
                double d2y = ny > 1 ? (mjp[k] - 2.0 * mij[k] + mjm[k]) / dy2 : 0.0;
                b[3 * c + k] += coeff * (d2x + d2y);
            }
        }
    }
}

/* dmi.rs:67-209.  Domain edge: chiral ghost (always); masked-out neighbour:
 * centre value (free boundary, NOT chiral); masked-out centre: skipped.    */
void orc_add_dmi(const orc_grid *g, const double *m, double *b,
                 const orc_material *mat, const uint8_t *mask) {
    if (!mat->has_dmi || mat->dmi == 0.0) return;
    double d = mat->dmi, ms = mat->ms, a = mat->a_ex;
    if (ms == 0.0 || a == 0.0) return;
    int64_t nx = g->nx, ny = g->ny;
    if (nx == 0 || ny == 0) return;
    double dx = g->dx, dy = g->dy;
    double pref = 2.0 * d / ms;
    double eta = d / (2.0 * a);

    for (int64_t j = 0; j < ny; ++j) {
        for (int64_t i = 0; i < nx; ++i) {
            int64_t c = j * nx + i;
            if (mask && !mask[c]) continue;
            const double *mij = m + 3 * c;
            double mim[3], mip[3], mjm[3], mjp[3];
            if (nx == 1) { copy3(mij, mim); copy3(mij, mip); }
            else {
                if (i == 0) ghost_x(mij, -1.0, eta, dx, mim);
                else { int64_t l = c - 1; if (!mask || mask[l]) copy3(m + 3 * l, mim); else copy3(mij, mim); }
                if (i == nx - 1) ghost_x(mij, 1.0, eta, dx, mip);
                else { int64_t r = c + 1; if (!mask || mask[r]) copy3(m + 3 * r, mip); else copy3(mij, mip); }
            }
            if (ny == 1) { copy3(mij, mjm); copy3(mij, mjp); }
            else {
                if (j == 0) ghost_y(mij, -1.0, eta, dy, mjm);
                else { int64_t dn = c - nx; if (!mask || mask[dn]) copy3(m + 3 * dn, mjm); else copy3(mij, mjm); }
                if (j == ny - 1) ghost_y(mij, 1.0, eta, dy, mjp);
                else { int64_t up = c + nx; if (!mask || mask[up]) copy3(m + 3 * up, mjp); else copy3(mij, mjp); }
            }
            double dmz_dx = nx > 1 ? (mip[2] - mim[2]) / (2.0 * dx) : 0.0;
            double dmz_dy = ny > 1 ? (mjp[2] - mjm[2]) / (2.0 * dy) : 0.0;
            double dmx_dx = nx > 1 ? (mip[0] - mim[0]) / (2.0 * dx) : 0.0;
            double dmy_dy = ny > 1 ? (mjp[1] - mjm[1]) / (2.0 * dy) : 0.0;
            b[3 * c + 0] += pref * dmz_dx;
            b[3 * c + 1] += pref * dmz_dy;
            b[3 * c + 2] += -pref * (dmx_dx + dmy_dy);
        }
    }
}

/* llg.rs:20-32 */
static inline void llg_rhs1(const double *m, const double *b, double gamma, double alpha, double *k) {
    double denom = 1.0 + alpha * alpha;
    double pref = -gamma / denom;
    double mxb[3], mxmxb[3];
    cross3(m, b, mxb);
    cross3(m, mxb, mxmxb);
    k[0] = pref * (mxb[0] + alpha * mxmxb[0]);
    k[1] = pref * (mxb[1] + alpha * mxmxb[1]);
    k[2] = pref * (mxb[2] + alpha * mxmxb[2]);
}
void orc_llg_rhs(int64_t n, const double *m, const double *b, double gamma, double alpha, double *k) {
    for (int64_t i = 0; i < n; ++i) llg_rhs1(m + 3 * i, b + 3 * i, gamma, alpha, k + 3 * i);
}

/* ------------------------------------------------------------------------ */
/* FFT: Stockham autosort, any length (stands in for rustfft)                */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n;
    int nf;
    int fac[64];
    double *tw; /* exp(-2 pi i k / n), k < n, interleaved */
} fft_plan;

static void plan_init(fft_plan *p, int64_t n) {
    p->n = n;
    p->nf = 0;
    int64_t r = n;
    static const int pref[4] = {4, 2, 3, 5};
    for (int t = 0; t < 4; ++t)
        while (r > 1 && r % pref[t] == 0) { p->fac[p->nf++] = pref[t]; r /= pref[t]; }
    for (int64_t q = 7; r > 1; q += 2)
        while (r % q == 0) { p->fac[p->nf++] = (int)q; r /= q; }
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    for (int64_t k = 0; k < n; ++k) {
        double ang = -2.0 * PI_D * (double)k / (double)n;
        p->tw[2 * k] = cos(ang);
        p->tw[2 * k + 1] = sin(ang);
    }
}
static void plan_free(fft_plan *p) { free(p->tw); p->tw = NULL; }

/* One transform of length p->n on x (interleaved complex), work of same size. */
static void fft_exec(const fft_plan *p, double *x, double *work, int inverse) {
    int64_t N = p->n;
    if (N <= 1) return;
    double sgn = inverse ? -1.0 : 1.0; /* multiplies the imaginary part of tw */
    double *src = x, *dst = work;
    int64_t Ns = 1;
    double vbuf[2 * 64], wbuf[2 * 64];
    for (int f = 0; f < p->nf; ++f) {
        int R = p->fac[f];
        /* radix buffers: on the stack for R <= 64, on the heap for larger primes */
        double *v = vbuf, *w = wbuf, *heap = NULL;
        if (R > 64) {
            heap = (double *)malloc(sizeof(double) * 4 * (size_t)R);
            v = heap;
            w = heap + 2 * R;
        }
        int64_t stride = N / R;
        int64_t tstep0 = N / (Ns * R);
        for (int64_t j = 0; j < N / R; ++j) {
            int64_t k = j % Ns;
            int64_t tstep = k * tstep0;
            for (int r = 0; r < R; ++r) {
                double re = src[2 * (j + r * stride)], im = src[2 * (j + r * stride) + 1];
                int64_t ti = (r * tstep) % N;
                double c = p->tw[2 * ti], s = sgn * p->tw[2 * ti + 1];
                v[2 * r] = re * c - im * s;
                v[2 * r + 1] = re * s + im * c;
            }
            if (R == 2) {
                w[0] = v[0] + v[2]; w[1] = v[1] + v[3];
                w[2] = v[0] - v[2]; w[3] = v[1] - v[3];
            } else if (R == 4) {
                double a0r = v[0] + v[4], a0i = v[1] + v[5];
                double a1r = v[0] - v[4], a1i = v[1] - v[5];
                double a2r = v[2] + v[6], a2i = v[3] + v[7];
                double a3r = v[2] - v[6], a3i = v[3] - v[7];
                /* multiply a3 by -i (forward) or +i (inverse) */
                double b3r = inverse ? -a3i : a3i, b3i = inverse ? a3r : -a3r;
                w[0] = a0r + a2r; w[1] = a0i + a2i;
                w[2] = a1r + b3r; w[3] = a1i + b3i;
                w[4] = a0r - a2r; w[5] = a0i - a2i;
                w[6] = a1r - b3r; w[7] = a1i - b3i;
            } else {
                for (int q = 0; q < R; ++q) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; ++r) {
                        int64_t ti = ((int64_t)q * r % R) * (N / R);
                        double c = p->tw[2 * ti], s = sgn * p->tw[2 * ti + 1];
                        sr += v[2 * r] * c - v[2 * r + 1] * s;
                        si += v[2 * r] * s + v[2 * r + 1] * c;
                    }
                    w[2 * q] = sr; w[2 * q + 1] = si;
                }
            }
            int64_t idxD = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) {
                dst[2 * (idxD + r * Ns)] = w[2 * r];
                dst[2 * (idxD + r * Ns) + 1] = w[2 * r + 1];
            }
        }
        double *t = src; src = dst; dst = t;
        Ns *= R;
        free(heap);
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)N);
}

void orc_fft1d(double *data, int64_t n, int inverse) {
    fft_plan p;
    plan_init(&p, n);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    fft_exec(&p, data, work, inverse);
    free(work);
    plan_free(&p);
}

/* fft2_forward_in_place / fft2_inverse_in_place (demag_fft_uniform.rs:777-922):
 * rows in parallel, then columns through a transpose, inverse scaled 1/(nx ny). */
static void fft2_planned(const fft_plan *pxp, const fft_plan *pyp, double *data,
                         double *tmp, int64_t nx, int64_t ny, int inverse) {
    #pragma omp parallel
    {
        double *work = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny));
        #pragma omp for schedule(static)
        for (int64_t y = 0; y < ny; ++y) fft_exec(pxp, data + 2 * y * nx, work, inverse);
        #pragma omp for schedule(static)
        for (int64_t x = 0; x < nx; ++x)
            for (int64_t y = 0; y < ny; ++y) {
                tmp[2 * (x * ny + y)] = data[2 * (y * nx + x)];
                tmp[2 * (x * ny + y) + 1] = data[2 * (y * nx + x) + 1];
            }
        #pragma omp for schedule(static)
        for (int64_t x = 0; x < nx; ++x) fft_exec(pyp, tmp + 2 * x * ny, work, inverse);
        #pragma omp for schedule(static)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t x = 0; x < nx; ++x) {
                data[2 * (y * nx + x)] = tmp[2 * (x * ny + y)];
                data[2 * (y * nx + x) + 1] = tmp[2 * (x * ny + y) + 1];
            }
        if (inverse) {
            double scale = 1.0 / (double)(nx * ny);
            #pragma omp for schedule(static)
            for (int64_t i = 0; i < 2 * nx * ny; ++i) data[i] *= scale;
        }
        free(work);
    }
}

void orc_fft2(double *data, int64_t nx, int64_t ny, int inverse) {
    fft_plan px, py;
    plan_init(&px, nx);
    plan_init(&py, ny);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)(nx * ny));
    fft2_planned(&px, &py, data, tmp, nx, ny, inverse);
    free(tmp);
    plan_free(&px);
    plan_free(&py);
}

/* ------------------------------------------------------------------------ */
/* Demag kernel: MuMax-like face-charge quadrature                           */
/* ------------------------------------------------------------------------ */
/* delta_cell (demag_fft_uniform.rs:1246-1253) */
static inline double delta_cell(int64_t d) {
    double a = (double)(d < 0 ? -d : d);
    if (a > 0.0) a -= 1.0;
    return a;
}
/* Planted-mistake switch for the tests that show what pins the quadrature:
 * 0 = the reference's rule; 1 = source faces not doubled (no nv *= 2,
 * nw *= 2); 2 = counts truncated instead of rounded (no + 0.5 in mumax_n).
 * Only tests set it; every product comparison runs with 0. */
static int g_quad_variant = 0;
void orc_set_quad_variant(int v) { g_quad_variant = v; }

/* mumax_n (demag_fft_uniform.rs:1158-1161) */
static inline int64_t mumax_n(double x) {
    return (int64_t)floor(fmax(x, 1.0) + (g_quad_variant == 2 ? 0.0 : 0.5));
}

/* mumax_like_h_from_unit_m (demag_fft_uniform.rs:1139-1242): same point
 * placement, loop nesting (v, w, x, y, z) and summation order. */
void orc_h_from_unit_m(int source_axis, const double r_center[3], const double cell[3],
                       const int64_t disp[3], double accuracy, double h[3]) {
    int u = source_axis, v = (u + 1) % 3, w = (u + 2) % 3;
    double lmin = fmin(fmin(cell[0], cell[1]), cell[2]);
    double dx_min = delta_cell(disp[0]) * cell[0];
    double dy_min = delta_cell(disp[1]) * cell[1];
    double dz_min = delta_cell(disp[2]) * cell[2];
    double d = sqrt(dx_min * dx_min + dy_min * dy_min + dz_min * dz_min);
    if (d == 0.0) d = lmin;
    double max_size = d / accuracy;
    int64_t nx = mumax_n(cell[0] / max_size);
    int64_t ny = mumax_n(cell[1] / max_size);
    int64_t nz = mumax_n(cell[2] / max_size);
    int64_t nv = mumax_n(cell[v] / max_size);
    int64_t nw = mumax_n(cell[w] / max_size);
    if (g_quad_variant != 1) {
        nv *= 2;
        nw *= 2;
    }
    double scale = 1.0 / (double)(nv * nw * nx * ny * nz);
    double face_area = cell[v] * cell[w];
    double charge = face_area * scale;
    double pu1 = 0.5 * cell[u];
    double pu2 = -pu1;
    double pole[3] = {0.0, 0.0, 0.0};
    h[0] = h[1] = h[2] = 0.0;
    for (int64_t i = 0; i < nv; ++i) {
        double pv = -0.5 * cell[v] + cell[v] / (2.0 * (double)nv) + (double)i * (cell[v] / (double)nv);
        pole[v] = pv;
        for (int64_t j = 0; j < nw; ++j) {
            double pw = -0.5 * cell[w] + cell[w] / (2.0 * (double)nw) + (double)j * (cell[w] / (double)nw);
            pole[w] = pw;
            for (int64_t ax = 0; ax < nx; ++ax) {
                double rx = r_center[0] - 0.5 * cell[0] + cell[0] / (2.0 * (double)nx) + (double)ax * (cell[0] / (double)nx);
                for (int64_t ay = 0; ay < ny; ++ay) {
                    double ry = r_center[1] - 0.5 * cell[1] + cell[1] / (2.0 * (double)ny) + (double)ay * (cell[1] / (double)ny);
                    for (int64_t az = 0; az < nz; ++az) {
                        double rz = r_center[2] - 0.5 * cell[2] + cell[2] / (2.0 * (double)nz) + (double)az * (cell[2] / (double)nz);
                        pole[u] = pu1;
                        double r1x = rx - pole[0], r1y = ry - pole[1], r1z = rz - pole[2];
                        double r1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
                        double q1 = charge / (4.0 * PI_D * r1 * r1 * r1);
                        double hx1 = r1x * q1, hy1 = r1y * q1, hz1 = r1z * q1;
                        pole[u] = pu2;
                        double r2x = rx - pole[0], r2y = ry - pole[1], r2z = rz - pole[2];
                        double r2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
                        double q2 = -charge / (4.0 * PI_D * r2 * r2 * r2);
                        h[0] += hx1 + r2x * q2;
                        h[1] += hy1 + r2y * q2;
                        h[2] += hz1 + r2z * q2;
                    }
                }
            }
        }
    }
}

/* kernel_2d_components_mumax_like (demag_fft_uniform.rs:1076-1097) */
void orc_kernel_components(double dx, double dy, double dz, int64_t sx, int64_t sy,
                           double accuracy, double out4[4]) {
    double r_center[3] = {(double)sx * dx, (double)sy * dy, 0.0};
    double cell[3] = {dx, dy, dz};
    int64_t disp[3] = {sx, sy, 0};
    double hx[3], hy[3], hz[3];
    orc_h_from_unit_m(0, r_center, cell, disp, accuracy, hx);
    orc_h_from_unit_m(1, r_center, cell, disp, accuracy, hy);
    orc_h_from_unit_m(2, r_center, cell, disp, accuracy, hz);
    out4[0] = ORC_MU0 * hx[0];
    out4[1] = ORC_MU0 * hy[0]; /* Hx from My */
    out4[2] = ORC_MU0 * hy[1];
    out4[3] = ORC_MU0 * hz[2];
}

static inline int64_t wrap_index(int64_t d, int64_t n) {
    int64_t v = d % n;
    if (v < 0) v += n;
    return v;
}

/* build_kernel_realspace_mumax_2d (demag_fft_uniform.rs:1013-1072): base
 * quadrant, four-fold reflection with parity (xx, yy, zz even; xy odd-odd),
 * then kxy = 0 on both axes.  Arrays are real, px*py, zero-initialised here. */
void orc_build_kernel_realspace(const orc_grid *g, int64_t px, int64_t py,
                                double *kxx, double *kxy, double *kyy, double *kzz) {
    int64_t nx = g->nx, ny = g->ny;
    int64_t rx_max = nx - 1, ry_max = ny - 1;
    size_t np = (size_t)(px * py);
    memset(kxx, 0, sizeof(double) * np);
    memset(kxy, 0, sizeof(double) * np);
    memset(kyy, 0, sizeof(double) * np);
    memset(kzz, 0, sizeof(double) * np);
    int64_t nb = (rx_max + 1) * (ry_max + 1);
    double *base = (double *)malloc(sizeof(double) * 4 * (size_t)nb);
    /* Each displacement is independent; the reference evaluates them serially,
     * the values do not depend on the order. */
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < nb; ++t) {
        int64_t sy = t / (rx_max + 1), sx = t % (rx_max + 1);
        orc_kernel_components(g->dx, g->dy, g->dz, sx, sy, ORC_DEMAG_ACCURACY, base + 4 * t);
    }
    for (int64_t sy = 0; sy <= ry_max; ++sy) {
        for (int64_t sx = 0; sx <= rx_max; ++sx) {
            const double *k = base + 4 * (sy * (rx_max + 1) + sx);
            for (int a = 0; a < 2; ++a) {
                int64_t sxs = a == 0 ? 1 : -1;
                for (int bb = 0; bb < 2; ++bb) {
                    int64_t sys = bb == 0 ? 1 : -1;
                    int64_t ix = wrap_index(sxs * sx, px), iy = wrap_index(sys * sy, py);
                    int64_t idx = iy * px + ix;
                    kxx[idx] = k[0];
                    kxy[idx] = (double)sxs * (double)sys * k[1];
                    kyy[idx] = k[2];
                    kzz[idx] = k[3];
                }
            }
        }
    }
    for (int64_t sy = -ry_max; sy <= ry_max; ++sy) kxy[wrap_index(sy, py) * px + wrap_index(0, px)] = 0.0;
    for (int64_t sx = -rx_max; sx <= rx_max; ++sx) kxy[wrap_index(0, py) * px + wrap_index(sx, px)] = 0.0;
    free(base);
}

/* build_kernel_realspace_pbc (demag_fft_uniform.rs:941-1005): every
 * displacement |sx| <= nx*pbc_x - 1 (nx - 1 for an open axis), likewise y,
 * evaluated directly and accumulated into the wrapped bin (sy outer, sx inner,
 * as the reference loops); then kxy = 0 on the bin axes. */
void orc_build_kernel_realspace_pbc(const orc_grid *g, int64_t px, int64_t py, int64_t pbc_x,
                                    int64_t pbc_y, double *kxx, double *kxy, double *kyy,
                                    double *kzz) {
    const int64_t nx = g->nx, ny = g->ny;
    const int64_t sx_max = pbc_x == 0 ? nx - 1 : nx * pbc_x - 1;
    const int64_t sy_max = pbc_y == 0 ? ny - 1 : ny * pbc_y - 1;
    const int64_t wx = 2 * sx_max + 1, wy = 2 * sy_max + 1;
    size_t np = (size_t)(px * py);
    memset(kxx, 0, sizeof(double) * np);
    memset(kxy, 0, sizeof(double) * np);
    memset(kyy, 0, sizeof(double) * np);
    memset(kzz, 0, sizeof(double) * np);
    double *vals = (double *)malloc(sizeof(double) * 4 * (size_t)(wx * wy));
    /* the values are independent of evaluation order; accumulation below is serial */
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < wx * wy; ++t) {
        int64_t sy = t / wx - sy_max, sx = t % wx - sx_max;
        orc_kernel_components(g->dx, g->dy, g->dz, sx, sy, ORC_DEMAG_ACCURACY, vals + 4 * t);
    }
    for (int64_t sy = -sy_max; sy <= sy_max; ++sy) {
        int64_t iy = wrap_index(sy, py);
        for (int64_t sx = -sx_max; sx <= sx_max; ++sx) {
            int64_t ix = wrap_index(sx, px);
            int64_t idx = iy * px + ix;
            const double *k = vals + 4 * ((sy + sy_max) * wx + (sx + sx_max));
            kxx[idx] += k[0];
            kxy[idx] += k[1];
            kyy[idx] += k[2];
            kzz[idx] += k[3];
        }
    }
    for (int64_t sy = -sy_max; sy <= sy_max; ++sy) kxy[wrap_index(sy, py) * px + wrap_index(0, px)] = 0.0;
    for (int64_t sx = -sx_max; sx <= sx_max; ++sx) kxy[wrap_index(0, py) * px + wrap_index(sx, px)] = 0.0;
    free(vals);
}

/* ------------------------------------------------------------------------ */
/* Demag2D (demag_fft_uniform.rs:434-769)                                     */
/* ------------------------------------------------------------------------ */
struct orc_demag {
    orc_grid g;
    int64_t pbc_x, pbc_y;
    int64_t px, py;
    fft_plan planx, plany;
    double *kxx, *kxy, *kyy, *kzz;   /* k-space, complex */
    double *mx, *my, *mz, *bx, *by, *bz, *tmp;
};

orc_demag *orc_demag_new(const orc_grid *g) { return orc_demag_new_pbc(g, 0, 0); }

/* Demag2D::new (demag_fft_uniform.rs:470-634): a periodic axis is not padded. */
orc_demag *orc_demag_new_pbc(const orc_grid *g, int64_t pbc_x, int64_t pbc_y) {
    orc_demag *d = (orc_demag *)calloc(1, sizeof(orc_demag));
    d->g = *g;
    d->pbc_x = pbc_x;
    d->pbc_y = pbc_y;
    d->px = pbc_x > 0 ? g->nx : 2 * g->nx;
    d->py = pbc_y > 0 ? g->ny : 2 * g->ny;
    size_t np = (size_t)(d->px * d->py);
    plan_init(&d->planx, d->px);
    plan_init(&d->plany, d->py);
    double **bufs[11] = {&d->kxx, &d->kxy, &d->kyy, &d->kzz, &d->mx, &d->my, &d->mz,
                         &d->bx, &d->by, &d->bz, &d->tmp};
    for (int i = 0; i < 11; ++i) *bufs[i] = (double *)calloc(2 * np, sizeof(double));
    double *re = (double *)malloc(sizeof(double) * 4 * np);
    if (pbc_x == 0 && pbc_y == 0)
        orc_build_kernel_realspace(g, d->px, d->py, re, re + np, re + 2 * np, re + 3 * np);
    else
        orc_build_kernel_realspace_pbc(g, d->px, d->py, pbc_x, pbc_y, re, re + np, re + 2 * np,
                                       re + 3 * np);
    double *ks[4] = {d->kxx, d->kxy, d->kyy, d->kzz};
    for (int c = 0; c < 4; ++c) {
        for (size_t i = 0; i < np; ++i) { ks[c][2 * i] = re[c * np + i]; ks[c][2 * i + 1] = 0.0; }
        fft2_planned(&d->planx, &d->plany, ks[c], d->tmp, d->px, d->py, 0);
    }
    free(re);
    return d;
}

void orc_demag_free(orc_demag *d) {
    if (!d) return;
    plan_free(&d->planx);
    plan_free(&d->plany);
    free(d->kxx); free(d->kxy); free(d->kyy); free(d->kzz);
    free(d->mx); free(d->my); free(d->mz);
    free(d->bx); free(d->by); free(d->bz); free(d->tmp);
    free(d);
}

void orc_demag_dims(const orc_demag *d, int64_t *px, int64_t *py) { *px = d->px; *py = d->py; }
void orc_demag_kspace(const orc_demag *d, double *kxx, double *kxy, double *kyy, double *kzz) {
    size_t nb = sizeof(double) * 2 * (size_t)(d->px * d->py);
    memcpy(kxx, d->kxx, nb); memcpy(kxy, d->kxy, nb);
    memcpy(kyy, d->kyy, nb); memcpy(kzz, d->kzz, nb);
}

/* Demag2D::add_field (demag_fft_uniform.rs:636-769): accumulates into b. */
void orc_demag_add_field(orc_demag *d, const double *m, double *b, double ms) {
    int64_t nx = d->g.nx, ny = d->g.ny, px = d->px, py = d->py;
    int64_t np = px * py;
    memset(d->mx, 0, sizeof(double) * 2 * (size_t)np);
    memset(d->my, 0, sizeof(double) * 2 * (size_t)np);
    memset(d->mz, 0, sizeof(double) * 2 * (size_t)np);
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            const double *v = m + 3 * (j * nx + i);
            d->mx[2 * (j * px + i)] = ms * v[0];
            d->my[2 * (j * px + i)] = ms * v[1];
            d->mz[2 * (j * px + i)] = ms * v[2];
        }
    fft2_planned(&d->planx, &d->plany, d->mx, d->tmp, px, py, 0);
    fft2_planned(&d->planx, &d->plany, d->my, d->tmp, px, py, 0);
    fft2_planned(&d->planx, &d->plany, d->mz, d->tmp, px, py, 0);
    #pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < np; ++t) {
        double mxr = d->mx[2 * t], mxi = d->mx[2 * t + 1];
        double myr = d->my[2 * t], myi = d->my[2 * t + 1];
        double mzr = d->mz[2 * t], mzi = d->mz[2 * t + 1];
        double axr = d->kxx[2 * t], axi = d->kxx[2 * t + 1];
        double ayr = d->kxy[2 * t], ayi = d->kxy[2 * t + 1];
        double byr_ = d->kyy[2 * t], byi_ = d->kyy[2 * t + 1];
        double czr = d->kzz[2 * t], czi = d->kzz[2 * t + 1];
        /* bx = kxx*mx + kxy*my ; by = kxy*mx + kyy*my ; bz = kzz*mz */
        d->bx[2 * t] = (axr * mxr - axi * mxi) + (ayr * myr - ayi * myi);
        d->bx[2 * t + 1] = (axr * mxi + axi * mxr) + (ayr * myi + ayi * myr);
        d->by[2 * t] = (ayr * mxr - ayi * mxi) + (byr_ * myr - byi_ * myi);
        d->by[2 * t + 1] = (ayr * mxi + ayi * mxr) + (byr_ * myi + byi_ * myr);
        d->bz[2 * t] = czr * mzr - czi * mzi;
        d->bz[2 * t + 1] = czr * mzi + czi * mzr;
    }
    fft2_planned(&d->planx, &d->plany, d->bx, d->tmp, px, py, 1);
    fft2_planned(&d->planx, &d->plany, d->by, d->tmp, px, py, 1);
    fft2_planned(&d->planx, &d->plany, d->bz, d->tmp, px, py, 1);
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t s = j * px + i;
            double *o = b + 3 * (j * nx + i);
            o[0] += d->bx[2 * s];
            o[1] += d->by[2 * s];
            o[2] += d->bz[2 * s];
        }
}

/* ------------------------------------------------------------------------ */
/* Context: cached operator + RK45 scratch (llg.rs:126-189)                   */
/* ------------------------------------------------------------------------ */
struct orc_ctx {
    orc_grid g;
    orc_demag *demag;
    orc_demag *demag_pbc;   /* the last periodic operator (keyed by pbc_x, pbc_y) */
    orc_mg *mg;
    orc_mg_cfg mg_cfg;
    orc_hyb_cfg hyb;        /* hybrid MG correction (disabled by default) */
    double *dk;             /* cached delta kernel, built for dk_cfg / dk_hyb / dk_r */
    orc_mg_cfg dk_cfg;
    orc_hyb_cfg dk_hyb;
    int64_t dk_r;
    double *ms_[8];      /* m2..m7 at [2..7] */
    double *bs[8];       /* b1..b7 */
    double *ks[8];       /* k1..k7 */
    int fsal_valid;
    int last_mask;
    double last_b_ext[3];
    int64_t builds;
};

orc_ctx *orc_ctx_new(const orc_grid *g) {
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->g = *g;
    size_t n3 = (size_t)(3 * g->nx * g->ny);
    for (int s = 1; s <= 7; ++s) {
        c->bs[s] = (double *)calloc(n3, sizeof(double));
        c->ks[s] = (double *)calloc(n3, sizeof(double));
        if (s >= 2) c->ms_[s] = (double *)calloc(n3, sizeof(double));
    }
    c->last_mask = -1;
    orc_mg_cfg_default(&c->mg_cfg);
    orc_hyb_cfg_default(&c->hyb);
    c->dk_r = -1;
    return c;
}

void orc_ctx_free(orc_ctx *c) {
    if (!c) return;
    orc_demag_free(c->demag);
    orc_demag_free(c->demag_pbc);
    orc_mg_free(c->mg);
    free(c->dk);
    for (int s = 1; s <= 7; ++s) { free(c->bs[s]); free(c->ks[s]); free(c->ms_[s]); }
    free(c);
}

int64_t orc_ctx_field_builds(const orc_ctx *c) { return c->builds; }
void orc_ctx_grid(const orc_ctx *c, orc_grid *g) { *g = c->g; }

/* demag.rs:47-67, 109-142: FFT (demag_fft_uniform.rs:157-172, open BC), or
 * MG for demag_method mg / dst (add_demag_field_poisson_mg,
 * demag_poisson_mg.rs:3605-3635: cached solver, rebuilt when the structure
 * changes, runtime knobs applied per call). */
void orc_add_demag(orc_ctx *c, const double *m, double *b, const orc_material *mat) {
    if (!mat->demag) return;
    if (mat->demag_method == 1 || mat->demag_method == 2) {
        if (c->mg && !orc_mg_same_structure(c->mg, &c->g, &c->mg_cfg)) {
            orc_mg_free(c->mg);
            c->mg = NULL;
        }
        if (!c->mg) c->mg = orc_mg_new(&c->g, &c->mg_cfg);
        orc_mg_apply_cfg(c->mg, &c->mg_cfg);
        if (!orc_hyb_enabled(&c->hyb)) {
            orc_mg_add_field(c->mg, m, b, mat->ms);
            return;
        }
        /* DemagPoissonMGHybrid::add_field with ensure_delta_kernel (:2412-2466):
           the delta kernel is rebuilt when its key (MG settings, hybrid
           settings, effective radius) changes */
        const int64_t r = orc_hyb_effective_radius(&c->g, &c->hyb);
        if (r > 0) {
            orc_hyb_cfg hk = c->hyb;
            hk.bscale = 0.0;
            hk.bmargin = 0;
            if (!c->dk || c->dk_r != r || memcmp(&c->dk_cfg, &c->mg_cfg, sizeof c->mg_cfg) != 0 ||
                memcmp(&c->dk_hyb, &hk, sizeof hk) != 0) {
                free(c->dk);
                c->dk = (double *)malloc(sizeof(double) * 4 * (size_t)((2 * r + 1) * (2 * r + 1)));
                orc_hyb_delta_kernel(&c->g, &c->mg_cfg, &c->hyb, r, mat->ms, c->dk);
                c->dk_cfg = c->mg_cfg;
                c->dk_hyb = hk;
                c->dk_r = r;
            }
        }
        orc_hyb_add_field(c->mg, &c->hyb, r, r > 0 ? c->dk : NULL, m, b, mat->ms, &c->g);
        return;
    }
    if (!c->demag) c->demag = orc_demag_new(&c->g);
    orc_demag_add_field(c->demag, m, b, mat->ms);
}

/* add_demag_field_pbc (demag.rs:70-143): MG and DST fall back to FFT when
 * either axis is periodic. */
void orc_add_demag_pbc(orc_ctx *c, const double *m, double *b, const orc_material *mat,
                       int64_t pbc_x, int64_t pbc_y) {
    if (!mat->demag) return;
    if (pbc_x == 0 && pbc_y == 0) {
        orc_add_demag(c, m, b, mat);
        return;
    }
    if (c->demag_pbc && (c->demag_pbc->pbc_x != pbc_x || c->demag_pbc->pbc_y != pbc_y)) {
        orc_demag_free(c->demag_pbc);
        c->demag_pbc = NULL;
    }
    if (!c->demag_pbc) c->demag_pbc = orc_demag_new_pbc(&c->g, pbc_x, pbc_y);
    orc_demag_add_field(c->demag_pbc, m, b, mat->ms);
}

void orc_ctx_set_mg_cfg(orc_ctx *c, const orc_mg_cfg *cfg) { c->mg_cfg = *cfg; }
orc_mg *orc_ctx_mg(orc_ctx *c) { return c->mg; }

/* build_h_eff_masked (mod.rs:48-75) / build_h_eff_masked_geom (mod.rs:78-101) */
void orc_build_h_eff(orc_ctx *c, const double *m, double *b, const orc_params *p,
                     const orc_material *mat, int field_mask, int use_geom,
                     const uint8_t *geom) {
    int64_t n = c->g.nx * c->g.ny;
    const uint8_t *gm = use_geom ? geom : NULL;
    memset(b, 0, sizeof(double) * 3 * (size_t)n);
    orc_add_zeeman(n, b, p->b_ext);
    orc_add_exchange(&c->g, m, b, mat, gm);
    orc_add_anisotropy(n, m, b, mat);
    int include_dmi = field_mask == ORC_EXCH_ANIS_DMI || field_mask == ORC_FULL;
    if (include_dmi && mat->has_dmi) orc_add_dmi(&c->g, m, b, mat, gm);
    if (field_mask == ORC_FULL && mat->demag) orc_add_demag(c, m, b, mat);
    c->builds++;
}

void orc_rk45_reset(orc_ctx *c) {
    c->fsal_valid = 0;
    c->last_mask = -1;
}

/* step_llg_rk45_recompute_field_masked_adaptive (llg.rs:795-1026) */
void orc_rk45_step(orc_ctx *c, double *m, orc_params *p, const orc_material *mat,
                   int field_mask, double max_err, double headroom, double dt_min,
                   double dt_max, double *eps_out, int *accepted_out, double *dt_used_out) {
    static const double A21 = 1.0 / 5.0;
    static const double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
    static const double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
    static const double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                        A54 = -212.0 / 729.0;
    static const double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                        A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
    static const double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                        B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
    static const double BH1 = 5179.0 / 57600.0, BH3 = 7571.0 / 16695.0, BH4 = 393.0 / 640.0,
                        BH5 = -92097.0 / 339200.0, BH6 = 187.0 / 2100.0, BH7 = 1.0 / 40.0;

    int64_t n = c->g.nx * c->g.ny;
    double gamma = p->gamma, alpha = p->alpha, dt0 = p->dt;
    double **K = c->ks, **Bf = c->bs, **M = c->ms_;

    int fsal_ok = c->fsal_valid && c->last_mask == field_mask &&
                  c->last_b_ext[0] == p->b_ext[0] && c->last_b_ext[1] == p->b_ext[1] &&
                  c->last_b_ext[2] == p->b_ext[2];
    if (fsal_ok) {
        memcpy(Bf[1], Bf[7], sizeof(double) * 3 * (size_t)n);
        memcpy(K[1], K[7], sizeof(double) * 3 * (size_t)n);
    } else {
        orc_build_h_eff(c, m, Bf[1], p, mat, field_mask, 0, NULL);
        orc_llg_rhs(n, m, Bf[1], gamma, alpha, K[1]);
    }

    /* Stage s state: v = m; v = v + (dt0*a_sj) k_j sequentially; normalise
     * (add_scaled, llg.rs:53-55, applied in order j = 1..s-1). */
    const double coef[7][6] = {
        {0}, {0},
        {A21},
        {A31, A32},
        {A41, A42, A43},
        {A51, A52, A53, A54},
        {A61, A62, A63, A64, A65},
    };
    for (int s = 2; s <= 7; ++s) {
        double *ms = M[s];
        for (int64_t i = 0; i < n; ++i) {
            double v[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
            if (s < 7) {
                for (int jj = 1; jj < s; ++jj) {
                    double sc = dt0 * coef[s][jj - 1];
                    const double *kj = K[jj] + 3 * i;
                    v[0] = v[0] + sc * kj[0];
                    v[1] = v[1] + sc * kj[1];
                    v[2] = v[2] + sc * kj[2];
                }
            } else {
                const double bw[5] = {B1, B3, B4, B5, B6};
                const int kk[5] = {1, 3, 4, 5, 6};
                for (int q = 0; q < 5; ++q) {
                    double sc = dt0 * bw[q];
                    const double *kj = K[kk[q]] + 3 * i;
                    v[0] = v[0] + sc * kj[0];
                    v[1] = v[1] + sc * kj[1];
                    v[2] = v[2] + sc * kj[2];
                }
            }
            normalize3(v, ms + 3 * i);
        }
        orc_build_h_eff(c, ms, Bf[s], p, mat, field_mask, 0, NULL);
        orc_llg_rhs(n, ms, Bf[s], gamma, alpha, K[s]);
    }

    /* Error estimate (llg.rs:948-997) */
    double err_inf = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double dmax = 0.0;
        for (int cc = 0; cc < 3; ++cc) {
            double th = B1 * K[1][3 * i + cc] + B3 * K[3][3 * i + cc] + B4 * K[4][3 * i + cc] +
                        B5 * K[5][3 * i + cc] + B6 * K[6][3 * i + cc];
            double tl = BH1 * K[1][3 * i + cc] + BH3 * K[3][3 * i + cc] + BH4 * K[4][3 * i + cc] +
                        BH5 * K[5][3 * i + cc] + BH6 * K[6][3 * i + cc] + BH7 * K[7][3 * i + cc];
            double dd = fabs(th - tl);
            if (cc == 0) dmax = dd; else dmax = fmax(dmax, dd);
        }
        if (dmax > err_inf) err_inf = dmax;
    }
    double eps = err_inf * dt0;

    /* Controller (llg.rs:999-1025) */
    double dt_next = eps > 0.0 ? dt0 * headroom * pow(max_err / eps, 0.2) : dt0 * 2.0;
    if (dt_next < dt_min) dt_next = dt_min;
    if (dt_next > dt_max) dt_next = dt_max;
    int accept = (eps <= max_err) || (dt0 <= dt_min * 1.0000000001);
    if (accept) {
        memcpy(m, M[7], sizeof(double) * 3 * (size_t)n);
        c->fsal_valid = 1;
        c->last_mask = field_mask;
        c->last_b_ext[0] = p->b_ext[0]; c->last_b_ext[1] = p->b_ext[1]; c->last_b_ext[2] = p->b_ext[2];
    } else {
        c->fsal_valid = 0;
        c->last_mask = -1;
    }
    p->dt = dt_next;
    *eps_out = eps;
    *accepted_out = accept;
    *dt_used_out = dt0;
}

void orc_ctx_set_hyb_cfg(orc_ctx *c, const orc_hyb_cfg *h) { c->hyb = *h; }

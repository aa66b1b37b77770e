// Per-cell local effective field (Zeeman, exchange, uniaxial anisotropy,
// interfacial DMI, optional demag addend, optional LLG torque and the next
// RK stage state formed from it), shared by the
// stand-alone local-field kernel and the kernel that fuses it with the
// inverse row FFT of the demag field.
//
// Terms follow llg-sim's effective_field/ (zeeman.rs:6-11,
// exchange.rs:45-222, anisotropy.rs:10-25, dmi.rs:67-209), summed in the
// order of build_h_eff_masked (mod.rs:48-75); the torque is llg_rhs
// (llg.rs:20-32) or, for relaxation, llg_rhs_relax (llg.rs:38-50).
#pragma once

#include "internal.hpp"

// Every function here is compiled without implicit multiply-add contraction:
// a sum of two products (a cross product, m.u, the Laplacian's two axes) could
// be fused either way, and the compiler's choice may differ between the
// kernels that inline this code (stand-alone, fused with the inverse row FFT,
// batched).  The multiply-adds are written out instead (mad: fused in FP32,
// where the FMA halves the arithmetic; a rounded product and sum in FP64, the
// reference's own rounding), so every kernel rounds the same way and the
// stepper gives bitwise the same result through any of them.

namespace llgmi {

// a * b + c
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) {
#pragma clang fp contract(off)
    return a * b + c;
}

template <typename T>
__device__ __forceinline__ void ghost_x(const T *m, T n_x, T eta, T dx, T *o) {
#pragma clang fp contract(off)
    o[0] = mad(n_x * dx, -eta * m[2], m[0]);
    o[1] = m[1];
    o[2] = mad(n_x * dx, eta * m[0], m[2]);
}
template <typename T>
__device__ __forceinline__ void ghost_y(const T *m, T n_y, T eta, T dy, T *o) {
#pragma clang fp contract(off)
    o[0] = m[0];
    o[1] = mad(n_y * dy, -eta * m[2], m[1]);
    o[2] = mad(n_y * dy, eta * m[1], m[2]);
}

// The torque of cell c from its m and total field B: llg_rhs (llg.rs:20-32),
// -gamma/(1+alpha^2) [m x B + alpha m x (m x B)], or for relaxation
// llg_rhs_relax (llg.rs:38-50), -gamma alpha/(1+alpha^2) m x (m x B).
template <typename T>
__device__ __forceinline__ void torque_cell(const LocalArgs<T> &a, const T *m0, const T *bv, T *kv) {
#pragma clang fp contract(off)
    T cx = mad(m0[1], bv[2], -(m0[2] * bv[1]));
    T cy = mad(m0[2], bv[0], -(m0[0] * bv[2]));
    T cz = mad(m0[0], bv[1], -(m0[1] * bv[0]));
    T ccx = mad(m0[1], cz, -(m0[2] * cy));
    T ccy = mad(m0[2], cx, -(m0[0] * cz));
    T ccz = mad(m0[0], cy, -(m0[1] * cx));
    if (a.torque == 2) {
        kv[0] = a.rhs_pref * ccx;
        kv[1] = a.rhs_pref * ccy;
        kv[2] = a.rhs_pref * ccz;
    } else {
        kv[0] = a.rhs_pref * mad(a.alpha, ccx, cx);
        kv[1] = a.rhs_pref * mad(a.alpha, ccy, cy);
        kv[2] = a.rhs_pref * mad(a.alpha, ccz, cz);
    }
}

// v = v + c_q k_q for q = 0 .. nk - 1 in order, then normalised (zero -> +z):
// the stage state of step_llg_rk45_recompute_field_masked_adaptive
// (llg.rs:870-941).  k_q is kp[q][o + comp * n] for q < nk - 1; the last term
// takes klast (when non-null) instead of memory.
template <typename T>
__device__ __forceinline__ void stage_combine(T *v, const T *const *kp, const T *c, int nk, const T *klast,
                                              int64_t o, int64_t n) {
#pragma clang fp contract(off)
    for (int q = 0; q < nk; ++q) {
        if (klast && q == nk - 1) {
            v[0] = mad(c[q], klast[0], v[0]);
            v[1] = mad(c[q], klast[1], v[1]);
            v[2] = mad(c[q], klast[2], v[2]);
        } else {
            const T *kq = kp[q];
            v[0] = mad(c[q], kq[o], v[0]);
            v[1] = mad(c[q], kq[n + o], v[1]);
            v[2] = mad(c[q], kq[2 * n + o], v[2]);
        }
    }
    T n2 = mad(v[2], v[2], mad(v[1], v[1], v[0] * v[0]));
    if (n2 == T(0)) {
        v[0] = T(0); v[1] = T(0); v[2] = T(1);
    } else {
        T inv = T(1) / sqrt(n2);
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    }
}

// Cell (i, j) of an nx x ny SoA field; bd3 = the cell's demag field (DEMAG).
template <typename T, bool GEOM, bool DEMAG, bool TORQUE>
__device__ __forceinline__ void local_cell(const LocalArgs<T> &a, const T *__restrict__ m,
                                           const uint8_t *__restrict__ geom, int i, int j, const T *bd3,
                                           T *__restrict__ b, T *__restrict__ kout) {
#pragma clang fp contract(off)
    const int nx = a.nx, ny = a.ny;
    const int64_t n = (int64_t)nx * ny;
    const int64_t c = (int64_t)j * nx + i;
    const T *mx = m, *my = m + n, *mz = m + 2 * n;
    T m0[3] = {mx[c], my[c], mz[c]};
    T bv[3] = {a.bext[0], a.bext[1], a.bext[2]};  // b = 0 + B_ext
    const bool inside = !GEOM || geom[c] != 0;

    if ((a.exch_on || a.dmi_on) && inside) {
        // raw neighbours (valid = exists in the domain and is material)
        T nl[3], nr[3], nd[3], nu[3];
        bool vl = false, vr = false, vd = false, vu = false;
        bool el = nx > 1 && i == 0, er = nx > 1 && i == nx - 1;   // domain edges
        bool ed = ny > 1 && j == 0, eu = ny > 1 && j == ny - 1;
        if (nx > 1 && i > 0) {
            int64_t q = c - 1;
            vl = !GEOM || geom[q] != 0;
            if (vl) { nl[0] = mx[q]; nl[1] = my[q]; nl[2] = mz[q]; }
        }
        if (nx > 1 && i < nx - 1) {
            int64_t q = c + 1;
            vr = !GEOM || geom[q] != 0;
            if (vr) { nr[0] = mx[q]; nr[1] = my[q]; nr[2] = mz[q]; }
        }
        if (ny > 1 && j > 0) {
            int64_t q = c - nx;
            vd = !GEOM || geom[q] != 0;
            if (vd) { nd[0] = mx[q]; nd[1] = my[q]; nd[2] = mz[q]; }
        }
        if (ny > 1 && j < ny - 1) {
            int64_t q = c + nx;
            vu = !GEOM || geom[q] != 0;
            if (vu) { nu[0] = mx[q]; nu[1] = my[q]; nu[2] = mz[q]; }
        }
        if (a.exch_on) {
            // exchange.rs:96-204: domain edge or masked neighbour -> chiral ghost
            // when D != 0, else the centre value; nx == 1 -> centre value.
            T xm[3], xp[3], ym[3], yp[3];
            if (vl) { xm[0] = nl[0]; xm[1] = nl[1]; xm[2] = nl[2]; }
            else if (a.has_eta && nx > 1) ghost_x(m0, T(-1), a.eta, a.dx, xm);
            else { xm[0] = m0[0]; xm[1] = m0[1]; xm[2] = m0[2]; }
            if (vr) { xp[0] = nr[0]; xp[1] = nr[1]; xp[2] = nr[2]; }
            else if (a.has_eta && nx > 1) ghost_x(m0, T(1), a.eta, a.dx, xp);
            else { xp[0] = m0[0]; xp[1] = m0[1]; xp[2] = m0[2]; }
            if (vd) { ym[0] = nd[0]; ym[1] = nd[1]; ym[2] = nd[2]; }
            else if (a.has_eta && ny > 1) ghost_y(m0, T(-1), a.eta, a.dy, ym);
            else { ym[0] = m0[0]; ym[1] = m0[1]; ym[2] = m0[2]; }
            if (vu) { yp[0] = nu[0]; yp[1] = nu[1]; yp[2] = nu[2]; }
            else if (a.has_eta && ny > 1) ghost_y(m0, T(1), a.eta, a.dy, yp);
            else { yp[0] = m0[0]; yp[1] = m0[1]; yp[2] = m0[2]; }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                // (xp - 2 m + xm) / dx^2 + (yp - 2 m + ym) / dy^2 (2 m is exact,
                // so the multiply-add rounds xp - 2 m once either way)
                const T lx = nx > 1 ? mad(T(-2), m0[q], xp[q]) + xm[q] : T(0);
                const T ly = ny > 1 ? mad(T(-2), m0[q], yp[q]) + ym[q] : T(0);
                bv[q] = mad(a.exch_coeff, mad(lx, a.inv_dx2, ly * a.inv_dy2), bv[q]);
            }
        }
        if (a.anis_on) {
            const T s = a.anis_coeff * mad(m0[2], a.u[2], mad(m0[1], a.u[1], m0[0] * a.u[0]));
#pragma unroll
            for (int q = 0; q < 3; ++q) bv[q] = mad(s, a.u[q], bv[q]);
        }
        if (a.dmi_on) {
            // dmi.rs:123-181: domain edge -> chiral ghost; masked neighbour -> centre.
            T xm[3], xp[3], ym[3], yp[3];
            if (el) ghost_x(m0, T(-1), a.eta, a.dx, xm);
            else if (vl) { xm[0] = nl[0]; xm[1] = nl[1]; xm[2] = nl[2]; }
            else { xm[0] = m0[0]; xm[1] = m0[1]; xm[2] = m0[2]; }
            if (er) ghost_x(m0, T(1), a.eta, a.dx, xp);
            else if (vr) { xp[0] = nr[0]; xp[1] = nr[1]; xp[2] = nr[2]; }
            else { xp[0] = m0[0]; xp[1] = m0[1]; xp[2] = m0[2]; }
            if (ed) ghost_y(m0, T(-1), a.eta, a.dy, ym);
            else if (vd) { ym[0] = nd[0]; ym[1] = nd[1]; ym[2] = nd[2]; }
            else { ym[0] = m0[0]; ym[1] = m0[1]; ym[2] = m0[2]; }
            if (eu) ghost_y(m0, T(1), a.eta, a.dy, yp);
            else if (vu) { yp[0] = nu[0]; yp[1] = nu[1]; yp[2] = nu[2]; }
            else { yp[0] = m0[0]; yp[1] = m0[1]; yp[2] = m0[2]; }
            T dmz_dx = nx > 1 ? (xp[2] - xm[2]) * a.inv_2dx : T(0);
            T dmz_dy = ny > 1 ? (yp[2] - ym[2]) * a.inv_2dy : T(0);
            T dmy_dy = ny > 1 ? (yp[1] - ym[1]) * a.inv_2dy : T(0);
            T div = nx > 1 ? mad(xp[0] - xm[0], a.inv_2dx, dmy_dy) : dmy_dy;   // dmx/dx + dmy/dy
            bv[0] = mad(a.dmi_pref, dmz_dx, bv[0]);
            bv[1] = mad(a.dmi_pref, dmz_dy, bv[1]);
            bv[2] = mad(-a.dmi_pref, div, bv[2]);
        }
    } else if (a.anis_on) {
        // anisotropy is not masked (mod.rs:91)
        const T s = a.anis_coeff * mad(m0[2], a.u[2], mad(m0[1], a.u[1], m0[0] * a.u[0]));
#pragma unroll
        for (int q = 0; q < 3; ++q) bv[q] = mad(s, a.u[q], bv[q]);
    }
    if (DEMAG) {
        bv[0] += bd3[0];
        bv[1] += bd3[1];
        bv[2] += bd3[2];
    }
    if (b) {   // null: the field itself is not wanted (RK stages keep only the torque)
        b[c] = bv[0];
        b[n + c] = bv[1];
        b[2 * n + c] = bv[2];
    }
    if (TORQUE) {
        T kv[3];
        torque_cell(a, m0, bv, kv);
        kout[c] = kv[0];
        kout[n + c] = kv[1];
        kout[2 * n + c] = kv[2];
        if (a.st_out) {
            T co[6];
            for (int q = 0; q < a.st_nk; ++q) co[q] = a.st_dt ? T(*a.st_dt * a.st_a[q]) : a.st_c[q];
            T v[3] = {a.st_m[c], a.st_m[n + c], a.st_m[2 * n + c]};
            stage_combine<T>(v, a.st_k, co, a.st_nk, kv, c, n);
            a.st_out[c] = v[0];
            a.st_out[n + c] = v[1];
            a.st_out[2 * n + c] = v[2];
        }
    }
}

}  // namespace llgmi

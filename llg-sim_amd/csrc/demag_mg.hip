// Multigrid Poisson demag on the GPU (FP64).
//
// Reference: demag_poisson_mg.rs (config :115-267, stencils :840-1189, levels
// :1425-1577, RHS :1595-1802, boundary :1804-1833, V-cycle :1877-1945, solve
// :1947-2057, gradient :2063-2098), mg_kernels.rs (smoothers, residual,
// transfers, gradient), mg_treecode.rs (charges, Barnes-Hut octree, dipole).
//
// Every kernel evaluates the reference's expressions in the reference's order
// with floating-point contraction off, so a solve reproduces the CPU
// restatement to the last bit wherever the reference's order is sequential
// per element (smoothers, residuals, transfers, RHS, gradient, treecode
// moments and evaluation).  Two reductions differ in order only: the
// residual max-norm (max is exact in any order) and the dipole moments of the
// DirichletDipole boundary (summed in a fixed parallel order).
//
// Layout: one FP64 array per level and quantity, (k*ny + j)*nx + i.  The
// residual is never stored: the restriction recomputes it per fine cell
// (each fine cell belongs to exactly one coarse cell), and the convergence
// check only needs its max-norm.
#pragma clang fp contract(off)

#include "mg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>

namespace llgmi {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0 = 4.0 * kPi * 1e-7;   // params.rs:15
constexpr int kMaxSt = 729;                  // Galerkin probe box 9^3
constexpr int kMaxDepth = 64;                // tree_max_depth clamp (from_env :259-261)
#ifdef LLGMI_MG_RED_BLOCKS
constexpr int kRedBlocks = LLGMI_MG_RED_BLOCKS;
#else
constexpr int kRedBlocks = 1024;             // blocks of the grid-stride reductions
#endif
constexpr int kNT = 256;

struct DevStencil {
    double center, diag;
    int n, pad_;
    double c[kMaxSt];
    int off[kMaxSt][3];
};

struct TNode {            // octree node in pre-order; leaf iff count > 0
    double cx, cy, cz, half;
    int skip, depth, start, count;
    int skip_depth, pad_;    // depth of node `skip` (0 past the end)
};

inline unsigned nblocks(int64_t n, int nt = kNT) {
    int64_t b = (n + nt - 1) / nt;
    if (b > 0x7fffffffLL) throw Error(3, "multigrid launch exceeds the grid limit");
    return (unsigned)(b > 0 ? b : 1);
}

__device__ __forceinline__ int64_t idx3(int i, int j, int k, int nx, int ny) {
    return ((int64_t)k * ny + j) * nx + i;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// offdiag_sum_at (mg_kernels.rs:72-94)
__device__ __forceinline__ double offdiag(const double *__restrict__ phi, int nx, int ny, int nz, int i,
                                          int j, int k, const DevStencil *__restrict__ st) {
    double sum = 0.0;
    const int n = st->n;
    for (int q = 0; q < n; ++q) {
        const int ii = clampi(i + st->off[q][0], 0, nx - 1);
        const int jj = clampi(j + st->off[q][1], 0, ny - 1);
        const int kk = clampi(k + st->off[q][2], 0, nz - 1);
        sum += st->c[q] * phi[idx3(ii, jj, kk, nx, ny)];
    }
    return sum;
}

// Off-diagonal sum of the 7-point stencil (offsets +x, -x, +y, -y, +z, -z, the
// order of Stencil3D::seven_point) at an interior cell: the clamp of
// offdiag_sum_at never acts there, so this is the same sum in the same order.
__device__ __forceinline__ double offdiag7(const double *__restrict__ phi, int64_t id, int nx, int64_t plane,
                                           const DevStencil *__restrict__ st) {
    double sum = 0.0;
    sum += st->c[0] * phi[id + 1];
    sum += st->c[1] * phi[id - 1];
    sum += st->c[2] * phi[id + nx];
    sum += st->c[3] * phi[id - nx];
    sum += st->c[4] * phi[id + plane];
    sum += st->c[5] * phi[id - plane];
    return sum;
}

// Cell t of the box boundary (z faces, then y faces, then x faces).
__device__ __forceinline__ void boundary_cell(int64_t t, int nx, int ny, int nz, int &i, int &j, int &k) {
    const int64_t a = (int64_t)nx * ny;
    if (t < 2 * a) {
        k = t < a ? 0 : nz - 1;
        t = t < a ? t : t - a;
        j = (int)(t / nx);
        i = (int)(t % nx);
        return;
    }
    t -= 2 * a;
    const int64_t b = (int64_t)nx * (nz - 2);
    if (t < 2 * b) {
        j = t < b ? 0 : ny - 1;
        t = t < b ? t : t - b;
        k = 1 + (int)(t / nx);
        i = (int)(t % nx);
        return;
    }
    t -= 2 * b;
    const int64_t c = (int64_t)(ny - 2) * (nz - 2);
    i = t < c ? 0 : nx - 1;
    t = t < c ? t : t - c;
    j = 1 + (int)(t % (ny - 2));
    k = 1 + (int)(t / (ny - 2));
}

__host__ __device__ inline int64_t boundary_count(int nx, int ny, int nz) {
    return 2 * (int64_t)nx * ny + 2 * (int64_t)nx * (nz - 2) + 2 * (int64_t)(ny - 2) * (nz - 2);
}

// ---------------------------------------------------------------------------
// RHS: build_rhs_from_m (demag_poisson_mg.rs:1595-1802) on the three planes
// k = offz-1, offz, offz+1 that can be nonzero (the rest of rhs stays 0).
// ---------------------------------------------------------------------------
struct RhsGeom {
    int px, py, pz, offx, offy, offz, nxm, nym;
    double dx, dy, dz, ms;
};

template <typename T>
__device__ __forceinline__ bool m_at(const T *__restrict__ m, int64_t nm, const RhsGeom &g, int pi, int pj,
                                     int pk, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (pi < 0 || pj < 0 || pk < 0) return false;
    if (pi >= g.px || pj >= g.py || pk >= g.pz) return false;
    if (pk != g.offz) return false;
    if (pi < g.offx || pi >= g.offx + g.nxm || pj < g.offy || pj >= g.offy + g.nym) return false;
    const int64_t id = (int64_t)(pj - g.offy) * g.nxm + (pi - g.offx);
    const double v0 = (double)m[id], v1 = (double)m[nm + id], v2 = (double)m[2 * nm + id];
    const double n2 = v0 * v0 + v1 * v1 + v2 * v2;
    if (n2 < 1e-30) return false;
    out[0] = g.ms * v0;
    out[1] = g.ms * v1;
    out[2] = g.ms * v2;
    return true;
}

__device__ __forceinline__ double face_val(bool ia, double a, bool ib, double b) {
    if (ia && ib) return 0.5 * (a + b);
    if (ia) return a;
    if (ib) return b;
    return 0.0;
}

template <typename T>
__global__ void __launch_bounds__(kNT) k_mg_rhs(const T *__restrict__ m, int64_t nm, RhsGeom g, int k_lo,
                                                int nk, double *__restrict__ rhs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)(g.px - 2) * (g.py - 2);
    if (t >= plane * nk) return;
    const int k = k_lo + (int)(t / plane);
    const int64_t r = t % plane;
    const int j = 1 + (int)(r / (g.px - 2));
    const int i = 1 + (int)(r % (g.px - 2));
    double c[3], xp[3], xm[3], yp[3], ym[3], zp[3], zm[3];
    const bool ci = m_at(m, nm, g, i, j, k, c);
    const bool xpi = m_at(m, nm, g, i + 1, j, k, xp);
    const bool xmi = m_at(m, nm, g, i - 1, j, k, xm);
    const bool ypi = m_at(m, nm, g, i, j + 1, k, yp);
    const bool ymi = m_at(m, nm, g, i, j - 1, k, ym);
    const bool zpi = m_at(m, nm, g, i, j, k + 1, zp);
    const bool zmi = m_at(m, nm, g, i, j, k - 1, zm);
    const double mx_p = face_val(ci, c[0], xpi, xp[0]);
    const double mx_m = face_val(xmi, xm[0], ci, c[0]);
    const double my_p = face_val(ci, c[1], ypi, yp[1]);
    const double my_m = face_val(ymi, ym[1], ci, c[1]);
    const double mz_p = face_val(ci, c[2], zpi, zp[2]);
    const double mz_m = face_val(zmi, zm[2], ci, c[2]);
    rhs[idx3(i, j, k, g.px, g.py)] = (mx_p - mx_m) / g.dx + (my_p - my_m) / g.dy + (mz_p - mz_m) / g.dz;
}

// ---------------------------------------------------------------------------
// Smoothers (mg_kernels.rs:121-240), residual (:249-285), transfers (:293-439)
// ---------------------------------------------------------------------------
// One colour of smooth_rb_sor, in place: a cell of one colour reads only
// cells of the other colour and itself, so the reference's tmp buffer is not
// needed.  Thread -> (interior row, h), i = first cell of the colour + 2h.
__global__ void __launch_bounds__(kNT) k_mg_rb(double *__restrict__ phi, const double *__restrict__ rhs,
                                               int nx, int ny, int nz, double sx, double sy, double sz,
                                               double denom, double omega, int color) {
    const int H = (nx - 1) / 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rows = (int64_t)(ny - 2) * (nz - 2);
    if (t >= rows * H) return;
    const int64_t row = t / H;
    const int h = (int)(t - row * H);
    const int j = 1 + (int)(row % (ny - 2));
    const int k = 1 + (int)(row / (ny - 2));
    const int i = 1 + ((1 + j + k + color) & 1) + 2 * h;
    if (i > nx - 2) return;
    const int64_t id = idx3(i, j, k, nx, ny);
    const int64_t plane = (int64_t)nx * ny;
    const double off = sx * (phi[id - 1] + phi[id + 1]) + sy * (phi[id - nx] + phi[id + nx]) +
                       sz * (phi[id - plane] + phi[id + plane]);
    const double pn = (off - rhs[id]) / denom;
    const double po = phi[id];
    phi[id] = po + omega * (pn - po);
}

// One full red-black SOR sweep (colour 0, then colour 1) from A into B.
// The colour-1 update reads the new colour-0 values and the old own value,
// the colour-0 update only old values, so with separate input and output
// buffers the sweep is the reference's in-place pair of half-sweeps exactly.
// A block owns an RBX x RBY tile of interior (i, j) columns and a chunk of
// RBZ planes and marches up k.  Two rings in LDS: the old values A of four
// planes on the tile plus a two-cell halo (each plane read from HBM once,
// when the march reaches it), and the new colour-0 values of planes k-1, k,
// k+1 on the tile plus a one-cell halo (halo cells and the plane below the
// chunk recomputed, boundary cells keep their value).  rbz = planes per
// chunk (z-chunks add blocks at the cost of recomputed planes).
constexpr int RBX = 32, RBY = 8, RBW = RBX + 2, RBH = RBY + 2, RAW = RBX + 4, RAH = RBY + 4;
__global__ void __launch_bounds__(RBX * RBY) k_mg_rb_sweep(const double *__restrict__ A, double *__restrict__ B,
                                                       const double *__restrict__ rhs, int nx, int ny,
                                                       int nz, double sx, double sy, double sz, double denom,
                                                       double omega, int rbz) {
    __shared__ double red[3][RBH][RBW];
    __shared__ double av[4][RAH][RAW];   // A on plane p in slot p % 4, origin (i0 - 2, j0 - 2)
    const int tid = threadIdx.x;
    const int i0 = 1 + (int)blockIdx.x * RBX, j0 = 1 + (int)blockIdx.y * RBY;
    const int kb = 1 + (int)blockIdx.z * rbz, ke = min(nz - 2, kb + rbz - 1);
    auto load = [&](int k) {
        if (k < 0 || k >= nz) return;
        double(*a)[RAW] = av[k & 3];
        for (int q = tid; q < RAH * RAW; q += RBX * RBY) {
            const int hj = q / RAW, hi = q - hj * RAW;
            const int i = i0 - 2 + hi, j = j0 - 2 + hj;
            a[hj][hi] = (i >= 0 && i < nx && j >= 0 && j < ny) ? A[idx3(i, j, k, nx, ny)] : 0.0;
        }
    };
    // new colour-0 value at an even cell (i, j, k), or the unchanged value of a
    // boundary cell; A from the LDS ring (planes k-1, k, k+1 loaded)
    auto red_new = [&](int i, int j, int k) -> double {
        const int hi = i - (i0 - 2), hj = j - (j0 - 2);
        const double(*ac)[RAW] = av[k & 3];
        if (i <= 0 || i >= nx - 1 || j <= 0 || j >= ny - 1 || k <= 0 || k >= nz - 1) return ac[hj][hi];
        const double(*am)[RAW] = av[(k - 1) & 3];
        const double(*ap)[RAW] = av[(k + 1) & 3];
        const double off = sx * (ac[hj][hi - 1] + ac[hj][hi + 1]) + sy * (ac[hj - 1][hi] + ac[hj + 1][hi]) +
                           sz * (am[hj][hi] + ap[hj][hi]);
        const double pn = (off - rhs[idx3(i, j, k, nx, ny)]) / denom;
        const double po = ac[hj][hi];
        return po + omega * (pn - po);
    };
    // only even-parity positions are ever read by the colour-1 update: the
    // in-plane and vertical neighbours of an odd cell are even
    auto fill = [&](int slot, int k) {
        for (int q = tid; q < (RBW / 2) * RBH; q += RBX * RBY) {
            const int hj = q / (RBW / 2);
            const int hi = 2 * (q % (RBW / 2)) + ((i0 + j0 + hj + k) & 1);
            const int i = i0 - 1 + hi, j = j0 - 1 + hj;
            red[slot][hj][hi] = (i < nx && j < ny) ? red_new(i, j, k) : 0.0;
        }
    };
    for (int k = kb - 2; k <= kb + 1; ++k) load(k);
    __syncthreads();
    fill((kb - 1) % 3, kb - 1);
    fill(kb % 3, kb);
    const int ti = tid % RBX, tj = tid / RBX;
    const int i = i0 + ti, j = j0 + tj;
    const bool inside = i <= nx - 2 && j <= ny - 2;
    // cells that never change: boundary columns and rows next to this tile
    // (the planes of this chunk, plus the bottom / top plane for the first /
    // last chunk) and the bottom / top planes under the tile
    if (inside) {
        const int klo = kb == 1 ? 0 : kb, khi = ke == nz - 2 ? nz - 1 : ke;
        auto keep = [&](int a, int c) {
            for (int k = klo; k <= khi; ++k) B[idx3(a, c, k, nx, ny)] = A[idx3(a, c, k, nx, ny)];
        };
        if (i == 1) keep(0, j);
        if (i == nx - 2) keep(nx - 1, j);
        if (j == 1) keep(i, 0);
        if (j == ny - 2) keep(i, ny - 1);
        if (i == 1 && j == 1) keep(0, 0);
        if (i == nx - 2 && j == 1) keep(nx - 1, 0);
        if (i == 1 && j == ny - 2) keep(0, ny - 1);
        if (i == nx - 2 && j == ny - 2) keep(nx - 1, ny - 1);
        if (kb == 1) B[idx3(i, j, 0, nx, ny)] = av[0][tj + 2][ti + 2];
        if (ke == nz - 2) B[idx3(i, j, nz - 1, nx, ny)] = A[idx3(i, j, nz - 1, nx, ny)];
    }
    for (int k = kb; k <= ke; ++k) {
        const int sm = (k - 1) % 3, sc = k % 3, sp = (k + 1) % 3;
        load(k + 2);   // slot of plane k - 2, which no step below reads
        __syncthreads();
        fill(sp, k + 1);
        __syncthreads();
        if (inside) {
            const int64_t id = idx3(i, j, k, nx, ny);
            const int hi = ti + 1, hj = tj + 1;
            if (((i + j + k) & 1) == 0) {
                B[id] = red[sc][hj][hi];
            } else {
                const double off = sx * (red[sc][hj][hi - 1] + red[sc][hj][hi + 1]) +
                                   sy * (red[sc][hj - 1][hi] + red[sc][hj + 1][hi]) +
                                   sz * (red[sm][hj][hi] + red[sp][hj][hi]);
                const double pn = (off - rhs[id]) / denom;
                const double po = av[k & 3][tj + 2][ti + 2];
                B[id] = po + omega * (pn - po);
            }
        }
    }
}

// Up to TBH red-black half-sweeps (colours alternating from `color`) in one
// launch, for small levels (the coarsest level's 80 sweeps).  A block loads
// an (TBT + 2 TBH)^2 x nz box of A and of rhs into LDS, updates the A box in
// place half-sweep by half-sweep (each half-sweep leaves one more outer ring
// stale), and writes the central TBT x TBT columns to B; after h <= TBH
// half-sweeps those are exactly the values the in-place half-sweeps produce.
// TBT = 10 puts the 131 x 131 coarsest level of the 512^2 film in 169 blocks
// (one per CU, one round on 256 CUs, 97 KB of LDS each); measured against 8,
// 12 and 16 it gives the shortest launches (fewer cells per block, no second
// round of blocks).
constexpr int TBT = 10, TBH = 8, TBW = TBT + 2 * TBH, TB_MAXZ = 9, TB_NT = 1024;
__device__ __forceinline__ int tb_at(int k, int hj, int hi) { return (k * TBW + hj) * TBW + hi; }
__global__ void __launch_bounds__(TB_NT) k_mg_rb_tb(const double *__restrict__ A, double *__restrict__ B,
                                                    const double *__restrict__ rhs, int nx, int ny, int nz,
                                                    double sx, double sy, double sz, double denom,
                                                    double omega, int color, int nhalf) {
    extern __shared__ unsigned char lds_raw[];
    const int nbox = nz * TBW * TBW;
    double *box = reinterpret_cast<double *>(lds_raw);   // A, [nz][TBW][TBW]
    double *rb = box + nbox;                               // rhs, same layout
    const int tid = threadIdx.x;
    const int i0 = 1 + (int)blockIdx.x * TBT, j0 = 1 + (int)blockIdx.y * TBT;
    const int ib = i0 - TBH, jb = j0 - TBH;   // global index of box column 0
    {
        // all global loads of this thread first, then the LDS stores
        constexpr int PER = (TB_MAXZ * TBW * TBW + TB_NT - 1) / TB_NT;
        double va[PER], vr[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * TB_NT;
            const int hi = q % TBW, hj = (q / TBW) % TBW, k = q / (TBW * TBW);
            const int gi = ib + hi, gj = jb + hj;
            const bool in = q < nbox && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
            va[u] = in ? A[idx3(gi, gj, k, nx, ny)] : 0.0;
            vr[u] = in ? rhs[idx3(gi, gj, k, nx, ny)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * TB_NT;
            if (q < nbox) {
                const int hi = q % TBW, hj = (q / TBW) % TBW, k = q / (TBW * TBW);
                box[tb_at(k, hj, hi)] = va[u];
                rb[tb_at(k, hj, hi)] = vr[u];
            }
        }
    }
    __syncthreads();
    constexpr int PER_ROW = (TBW - 2) / 2;   // cells of one colour per box row
    const int rows = (nz - 2) * (TBW - 2);
    for (int s = 0; s < nhalf; ++s) {
        const int c = (color + s) & 1;
        for (int q = tid; q < rows * PER_ROW; q += TB_NT) {
            const int row = q / PER_ROW, r = q % PER_ROW;
            const int hj = 1 + row % (TBW - 2), k = 1 + row / (TBW - 2);
            const int gj = jb + hj;
            const int first = 1 + (((ib + 1 + gj + k) & 1) == c ? 0 : 1);
            const int hi = first + 2 * r;
            const int gi = ib + hi;
            if (gi < 1 || gi > nx - 2 || gj < 1 || gj > ny - 2) continue;
            const int id = tb_at(k, hj, hi);
            const double off = sx * (box[tb_at(k, hj, hi - 1)] + box[tb_at(k, hj, hi + 1)]) +
                               sy * (box[tb_at(k, hj - 1, hi)] + box[tb_at(k, hj + 1, hi)]) +
                               sz * (box[tb_at(k - 1, hj, hi)] + box[tb_at(k + 1, hj, hi)]);
            const double pn = (off - rb[id]) / denom;
            const double po = box[id];
            box[id] = po + omega * (pn - po);
        }
        __syncthreads();
    }
    // central columns (and the untouched boundary cells next to them) to B
    for (int q = tid; q < nz * TBT * TBT; q += TB_NT) {
        const int ci = q % TBT, cj = (q / TBT) % TBT, k = q / (TBT * TBT);
        const int gi = i0 + ci, gj = j0 + cj;
        if (gi > nx - 2 || gj > ny - 2) continue;
        B[idx3(gi, gj, k, nx, ny)] = box[tb_at(k, cj + TBH, ci + TBH)];
        if (gi == 1) B[idx3(0, gj, k, nx, ny)] = A[idx3(0, gj, k, nx, ny)];
        if (gi == nx - 2) B[idx3(nx - 1, gj, k, nx, ny)] = A[idx3(nx - 1, gj, k, nx, ny)];
        if (gj == 1) B[idx3(gi, 0, k, nx, ny)] = A[idx3(gi, 0, k, nx, ny)];
        if (gj == ny - 2) B[idx3(gi, ny - 1, k, nx, ny)] = A[idx3(gi, ny - 1, k, nx, ny)];
        if (gi == 1 && gj == 1) B[idx3(0, 0, k, nx, ny)] = A[idx3(0, 0, k, nx, ny)];
        if (gi == nx - 2 && gj == 1) B[idx3(nx - 1, 0, k, nx, ny)] = A[idx3(nx - 1, 0, k, nx, ny)];
        if (gi == 1 && gj == ny - 2) B[idx3(0, ny - 1, k, nx, ny)] = A[idx3(0, ny - 1, k, nx, ny)];
        if (gi == nx - 2 && gj == ny - 2) B[idx3(nx - 1, ny - 1, k, nx, ny)] = A[idx3(nx - 1, ny - 1, k, nx, ny)];
    }
}

// One sweep of smooth_weighted_jacobi from src into dst; boundary cells of dst
// take the Dirichlet values (bc, or 0 when bc is null).
__global__ void __launch_bounds__(kNT) k_mg_wj(const double *__restrict__ src, double *__restrict__ dst,
                                               const double *__restrict__ rhs, const double *__restrict__ bc,
                                               const DevStencil *__restrict__ st, double omega, int nx,
                                               int ny, int nz) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nx * ny * nz) return;
    const int i = (int)(id % nx);
    const int64_t r = id / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1) {
        dst[id] = bc ? bc[id] : 0.0;
        return;
    }
    const double off = offdiag(src, nx, ny, nz, i, j, k, st);
    const double s = src[id];
    const double gs = fabs(st->diag) > 1e-30 ? (off - rhs[id]) / st->diag : s;
    dst[id] = (1.0 - omega) * s + omega * gs;
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[kNT / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmax(m, red[q]);
        v = m;
    }
    return v;
}

// max |rhs - A phi| over interior cells, per block (compute_residual's return value)
__global__ void __launch_bounds__(kNT) k_mg_resid_max(const double *__restrict__ phi,
                                                      const double *__restrict__ rhs,
                                                      const DevStencil *__restrict__ st, int nx, int ny,
                                                      int nz, int seven, double *__restrict__ partial) {
    const int64_t n = (int64_t)(nx - 2) * (ny - 2) * (nz - 2);
    const int64_t plane = (int64_t)nx * ny;
    double mx = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = 1 + (int)(t % (nx - 2));
        const int64_t r = t / (nx - 2);
        const int j = 1 + (int)(r % (ny - 2));
        const int k = 1 + (int)(r / (ny - 2));
        const int64_t id = idx3(i, j, k, nx, ny);
        const double a = st->center * phi[id] +
                         (seven ? offdiag7(phi, id, nx, plane, st) : offdiag(phi, nx, ny, nz, i, j, k, st));
        mx = fmax(mx, fabs(rhs[id] - a));
    }
    mx = block_max(mx);
    if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

// Residual + restrict_residual: coarse rhs = mean over the fine block of
// r = rhs - A phi; coarse phi = 0 (interior; the coarse boundary stays 0).
__global__ void __launch_bounds__(kNT) k_mg_resid_restrict(const double *__restrict__ phi,
                                                           const double *__restrict__ rhs,
                                                           const DevStencil *__restrict__ st, int fnx,
                                                           int fny, int fnz, double *__restrict__ crhs,
                                                           double *__restrict__ cphi, int cnx, int cny,
                                                           int cnz, int rx, int ry, int rz, double w,
                                                           int seven) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)(cnx - 2) * (cny - 2) * (cnz - 2);
    if (t >= n) return;
    const int ic = 1 + (int)(t % (cnx - 2));
    const int64_t r = t / (cnx - 2);
    const int jc = 1 + (int)(r % (cny - 2));
    const int kc = 1 + (int)(r / (cny - 2));
    double sum = 0.0;
    for (int dk = 0; dk < rz; ++dk)
        for (int dj = 0; dj < ry; ++dj)
            for (int di = 0; di < rx; ++di) {
                const int i = rx * ic + di, j = ry * jc + dj, k = rz * kc + dk;
                const int64_t id = idx3(i, j, k, fnx, fny);
                const double a = st->center * phi[id] +
                                 (seven ? offdiag7(phi, id, fnx, (int64_t)fnx * fny, st)
                                        : offdiag(phi, fnx, fny, fnz, i, j, k, st));
                sum += rhs[id] - a;
            }
    const int64_t cid = idx3(ic, jc, kc, cnx, cny);
    crhs[cid] = w * sum;
    cphi[cid] = 0.0;
}

__device__ __forceinline__ void interp_1d(int ic, int ri, int nc, int r, int &i0o, int &i1o, double &w0,
                                          double &w1) {
    const int i0 = ic < nc - 1 ? ic : nc - 1;
    if (r == 1) {
        i0o = i1o = i0;
        w0 = 1.0;
        w1 = 0.0;
        return;
    }
    const int i1 = ri == 0 ? (i0 > 0 ? i0 - 1 : 0) : (i0 + 1 < nc - 1 ? i0 + 1 : nc - 1);
    if (i1 == i0) {
        i0o = i1o = i0;
        w0 = 1.0;
        w1 = 0.0;
    } else {
        i0o = i0;
        i1o = i1;
        w0 = 0.75;
        w1 = 0.25;
    }
}

// prolongate_add (mg_kernels.rs:377-439), one thread per fine interior cell
__global__ void __launch_bounds__(kNT) k_mg_prolong(const double *__restrict__ c, int cnx, int cny, int cnz,
                                                    double *__restrict__ f, int fnx, int fny, int fnz,
                                                    int trilinear) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)(fnx - 2) * (fny - 2) * (fnz - 2);
    if (t >= n) return;
    const int i = 1 + (int)(t % (fnx - 2));
    const int64_t r = t / (fnx - 2);
    const int j = 1 + (int)(r % (fny - 2));
    const int k = 1 + (int)(r / (fny - 2));
    const int rx = fnx / cnx, ry = fny / cny, rz = fnz / cnz;
    const int64_t id = idx3(i, j, k, fnx, fny);
    if (trilinear) {
        int k0, k1, j0, j1, i0, i1;
        double wk0, wk1, wj0, wj1, wi0, wi1;
        interp_1d(k / rz, k % rz, cnz, rz, k0, k1, wk0, wk1);
        interp_1d(j / ry, j % ry, cny, ry, j0, j1, wj0, wj1);
        interp_1d(i / rx, i % rx, cnx, rx, i0, i1, wi0, wi1);
        const double v = wi0 * wj0 * wk0 * c[idx3(i0, j0, k0, cnx, cny)]
                       + wi1 * wj0 * wk0 * c[idx3(i1, j0, k0, cnx, cny)]
                       + wi0 * wj1 * wk0 * c[idx3(i0, j1, k0, cnx, cny)]
                       + wi1 * wj1 * wk0 * c[idx3(i1, j1, k0, cnx, cny)]
                       + wi0 * wj0 * wk1 * c[idx3(i0, j0, k1, cnx, cny)]
                       + wi1 * wj0 * wk1 * c[idx3(i1, j0, k1, cnx, cny)]
                       + wi0 * wj1 * wk1 * c[idx3(i0, j1, k1, cnx, cny)]
                       + wi1 * wj1 * wk1 * c[idx3(i1, j1, k1, cnx, cny)];
        f[id] += v;
    } else {
        const int ic = i / rx, jc = j / ry, kc = k / rz;
        if (ic == 0 || ic + 1 >= cnx || jc == 0 || jc + 1 >= cny || kc == 0 || kc + 1 >= cnz) return;
        f[id] += c[idx3(ic, jc, kc, cnx, cny)];
    }
}

// Dirichlet stamp of the boundary cells (stamp_dirichlet_bc, mg_kernels.rs:35-61)
__global__ void __launch_bounds__(kNT) k_mg_stamp(double *__restrict__ phi, const double *__restrict__ bc,
                                                  int nx, int ny, int nz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= boundary_count(nx, ny, nz)) return;
    int i, j, k;
    boundary_cell(t, nx, ny, nz, i, j, k);
    const int64_t id = idx3(i, j, k, nx, ny);
    phi[id] = bc ? bc[id] : 0.0;
}

// NaN/Inf guard of solve_with_timing (:2018-2033): the first 2 nx*ny values
__global__ void __launch_bounds__(kNT) k_mg_nan(const double *__restrict__ phi, int64_t n,
                                                double *__restrict__ flag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !isfinite(phi[t])) *flag = 1.0;
}

// max |rhs| over interior cells of the planes [k_lo, k_lo + nk)
__global__ void __launch_bounds__(kNT) k_mg_rhs_absmax(const double *__restrict__ rhs, int nx, int ny, int k_lo,
                                                       int nk, double *__restrict__ partial) {
    const int64_t plane = (int64_t)(nx - 2) * (ny - 2);
    const int64_t n = plane * nk;
    double mx = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int k = k_lo + (int)(t / plane);
        const int64_t r = t % plane;
        const int j = 1 + (int)(r / (nx - 2));
        const int i = 1 + (int)(r % (nx - 2));
        mx = fmax(mx, fabs(rhs[idx3(i, j, k, nx, ny)]));
    }
    mx = block_max(mx);
    if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

__global__ void __launch_bounds__(kNT) k_mg_final_max(const double *__restrict__ partial, int n,
                                                      double *__restrict__ out) {
    double mx = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) mx = fmax(mx, partial[q]);
    mx = block_max(mx);
    if (threadIdx.x == 0) *out = mx;
}

// extract_gradient_on_magnet_layer (mg_kernels.rs:450-507) as B_demag (overwrite)
template <typename T>
__global__ void __launch_bounds__(kNT) k_mg_gradient(const double *__restrict__ phi, int px, int py, int pz,
                                                     double dx, double dy, double dz, int offx, int offy,
                                                     int kmag, int nxm, int nym, const T *__restrict__ m,
                                                     T *__restrict__ bd) {
    const int64_t n = (int64_t)nxm * nym;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = (int)(id % nxm), j = (int)(id / nxm);
    if (m) {   // m null: every film cell (add_b_from_phi_on_magnet_layer_all)
        const double v0 = (double)m[id], v1 = (double)m[n + id], v2 = (double)m[2 * n + id];
        if (!(v0 * v0 + v1 * v1 + v2 * v2 > 1e-30)) {
            bd[id] = T(0);
            bd[n + id] = T(0);
            bd[2 * n + id] = T(0);
            return;
        }
    }
    const int pi = offx + i, pj = offy + j, k = kmag;
    const double pc = phi[idx3(pi, pj, k, px, py)];
    double gx = 0.0, wx = 0.0;
    if (pi + 1 < px) { gx += (phi[idx3(pi + 1, pj, k, px, py)] - pc) / dx; wx += 1.0; }
    if (pi > 0) { gx += (pc - phi[idx3(pi - 1, pj, k, px, py)]) / dx; wx += 1.0; }
    if (wx > 0.0) gx /= wx;
    double gy = 0.0, wy = 0.0;
    if (pj + 1 < py) { gy += (phi[idx3(pi, pj + 1, k, px, py)] - pc) / dy; wy += 1.0; }
    if (pj > 0) { gy += (pc - phi[idx3(pi, pj - 1, k, px, py)]) / dy; wy += 1.0; }
    if (wy > 0.0) gy /= wy;
    double gz = 0.0, wz = 0.0;
    if (k + 1 < pz) { gz += (phi[idx3(pi, pj, k + 1, px, py)] - pc) / dz; wz += 1.0; }
    if (k > 0) { gz += (pc - phi[idx3(pi, pj, k - 1, px, py)]) / dz; wz += 1.0; }
    if (wz > 0.0) gz /= wz;
    bd[id] = T(-kMu0 * gx);
    bd[n + id] = T(-kMu0 * gy);
    bd[2 * n + id] = T(-kMu0 * gz);
}

// screen_rhs_gaussian_xy (mg_kernels.rs:604-680): one pass of the separable
// Gaussian along x (AXIS 0: src = rhs, dst = tmp) or y (AXIS 1: src = tmp,
// dst = rhs); boundary rows and the row ends are zero, sources clamped.
template <int AXIS>
__global__ void __launch_bounds__(kNT) k_mg_screen(const double *__restrict__ src, double *__restrict__ dst,
                                                   const double *__restrict__ w, int r, int nx, int ny, int nz) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nx * ny * nz) return;
    const int64_t row = id / nx;
    const int i = (int)(id - row * nx);
    const int k = (int)(row / ny), j = (int)(row - (int64_t)k * ny);
    if (k == 0 || k + 1 == nz || j == 0 || j + 1 == ny || i == 0 || i == nx - 1) {
        dst[id] = 0.0;
        return;
    }
    double acc = 0.0;
    for (int q = -r; q <= r; ++q) {
        if (AXIS == 0) {
            int x = i + q;
            x = x < 0 ? 0 : (x > nx - 1 ? nx - 1 : x);
            acc += w[q + r] * src[row * nx + x];
        } else {
            int y = j + q;
            y = y < 0 ? 0 : (y > ny - 1 ? ny - 1 : y);
            acc += w[q + r] * src[((int64_t)k * ny + y) * nx + i];
        }
    }
    dst[id] = acc;
}

// Boundary scaling of the hybrid correction (demag_poisson_mg.rs:424-501):
// material cells within `margin` of vacuum (|m|^2 < 0.25) get
// (material fraction in the (2r+1)^2 window)^alpha, every other cell 1.
template <typename T>
__global__ void __launch_bounds__(kNT) k_hyb_scale(const T *__restrict__ m, int nx, int ny, int r, int margin,
                                                   double alpha, double *__restrict__ scale) {
    const int64_t n = (int64_t)nx * ny;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    auto vac = [&](int64_t q) {
        const double a = (double)m[q], b = (double)m[n + q], d = (double)m[2 * n + q];
        return (a * a + b * b + d * d) < 0.25;
    };
    const int i = (int)(c % nx), j = (int)(c / nx);
    double s = 1.0;
    if (!vac(c)) {
        bool near = false;
        for (int dj = -margin; dj <= margin && !near; ++dj)
            for (int di = -margin; di <= margin; ++di) {
                const int ni = i + di, nj = j + dj;
                if (ni < 0 || ni >= nx || nj < 0 || nj >= ny) continue;
                if (vac((int64_t)nj * nx + ni)) { near = true; break; }
            }
        if (near) {
            unsigned n_mat = 0, n_tot = 0;
            for (int dy = -r; dy <= r; ++dy) {
                const int sj = j + dy;
                if (sj < 0 || sj >= ny) continue;
                for (int dx = -r; dx <= r; ++dx) {
                    const int si = i + dx;
                    if (si < 0 || si >= nx) continue;
                    ++n_tot;
                    if (!vac((int64_t)sj * nx + si)) ++n_mat;
                }
            }
            s = pow((double)n_mat / (double)(n_tot > 1 ? n_tot : 1), alpha);
        }
    }
    scale[c] = s;
}

// DeltaKernel2D::add_correction (demag_poisson_mg.rs:502-583): b += dK * (Ms m)
// over the (2r+1)^2 stencil (sources outside the film skipped), summed in the
// reference's order; dk = [xx | xy | yy | zz].
template <typename T>
__global__ void __launch_bounds__(kNT) k_hyb_correct(const double *__restrict__ dk, int r, const T *__restrict__ m,
                                                     T *__restrict__ b, double ms, const double *__restrict__ scale,
                                                     int nx, int ny) {
    const int64_t n = (int64_t)nx * ny;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % nx), j = (int)(c / nx), st = 2 * r + 1, nst = st * st;
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int dy = -r; dy <= r; ++dy) {
        const int sj = j - dy;
        if (sj < 0 || sj >= ny) continue;
        for (int dx = -r; dx <= r; ++dx) {
            const int si = i - dx;
            if (si < 0 || si >= nx) continue;
            const int k = (dy + r) * st + (dx + r);
            const int64_t q = (int64_t)sj * nx + si;
            const double mx = ms * (double)m[q], my = ms * (double)m[n + q], mz = ms * (double)m[2 * n + q];
            bx += dk[k] * mx + dk[nst + k] * my;
            by += dk[nst + k] * mx + dk[2 * nst + k] * my;
            bz += dk[3 * nst + k] * mz;
        }
    }
    if (scale) {
        const double sc = scale[c];
        bx *= sc;
        by *= sc;
        bz *= sc;
    }
    b[c] = T((double)b[c] + bx);
    b[n + c] = T((double)b[n + c] + by);
    b[2 * n + c] = T((double)b[2 * n + c] + bz);
}

// ---------------------------------------------------------------------------
// Treecode (mg_treecode.rs): charges, moments, boundary evaluation
// ---------------------------------------------------------------------------
// Candidate cells: interior cells of the planes [k_lo, k_lo + nk) in raster
// order (the only cells the RHS can make nonzero).
__device__ __forceinline__ int64_t cand_cell(int64_t t, int nx, int ny, int k_lo) {
    const int64_t plane = (int64_t)(nx - 2) * (ny - 2);
    const int k = k_lo + (int)(t / plane);
    const int64_t r = t % plane;
    return idx3(1 + (int)(r % (nx - 2)), 1 + (int)(r / (nx - 2)), k, nx, ny);
}

__global__ void __launch_bounds__(kNT) k_tc_count(const double *__restrict__ rhs, int nx, int ny, int k_lo,
                                                  int64_t ncand, int *__restrict__ counts) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0;
    if (t < ncand) f = fabs(rhs[cand_cell(t, nx, ny, k_lo)]) < 1e-40 ? 0 : 1;
    const int c = __syncthreads_count(f);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// Exclusive scan of the block counts (one block); meta[1] = total.
__global__ void __launch_bounds__(1024) k_tc_scan(const int *__restrict__ counts, int nb, int *__restrict__ offsets,
                                                  int *__restrict__ meta) {
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = min(nb, lo + per);
    int s = 0;
    for (int q = lo; q < hi; ++q) s += counts[q];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int q = 0; q < 1024; ++q) {
            const int v = part[q];
            part[q] = acc;
            acc += v;
        }
        meta[1] = acc;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int q = lo; q < hi; ++q) {
        offsets[q] = acc;
        acc += counts[q];
    }
}

// Compacted charges in raster order: cell index and q = rhs dV
// (build_charges_from_rhs, mg_treecode.rs:273-300); meta[0] = 1 when the
// list differs from the cached one.
__global__ void __launch_bounds__(kNT) k_tc_write(const double *__restrict__ rhs, int nx, int ny, int k_lo,
                                                  int64_t ncand, const int *__restrict__ offsets, double dvol,
                                                  int *__restrict__ idx_out, double *__restrict__ q_out,
                                                  const int *__restrict__ cached, int cached_n,
                                                  int *__restrict__ meta) {
    __shared__ int wsum[kNT / 64];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t cell = 0;
    double rho = 0.0;
    bool f = false;
    if (t < ncand) {
        cell = cand_cell(t, nx, ny, k_lo);
        rho = rhs[cell];
        f = !(fabs(rho) < 1e-40);
    }
    const unsigned long long bal = __ballot(f);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int base = offsets[blockIdx.x];
    for (int q = 0; q < w; ++q) base += wsum[q];
    if (f) {
        const int pos = base + before;
        idx_out[pos] = (int)cell;
        q_out[pos] = rho * dvol;
        if (pos >= cached_n || cached[pos] != (int)cell) meta[0] = 1;
    }
}

__global__ void __launch_bounds__(kNT) k_tc_gather(const double *__restrict__ q, const int *__restrict__ ord, int n,
                                                   double *__restrict__ qp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) qp[t] = q[ord[t]];
}

// compute_moments_rec (mg_treecode.rs:158-190), leaves
__global__ void __launch_bounds__(kNT) k_tc_leaf_moments(const TNode *__restrict__ nodes,
                                                         const int *__restrict__ leaves, int nleaves,
                                                         const double *__restrict__ pos,
                                                         const double *__restrict__ qp,
                                                         double *__restrict__ mom) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nleaves) return;
    const int id = leaves[t];
    const TNode nd = nodes[id];
    double q = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int a = nd.start; a < nd.start + nd.count; ++a) {
        const double cq = qp[a];
        q += cq;
        p0 += cq * (pos[3 * (int64_t)a] - nd.cx);
        p1 += cq * (pos[3 * (int64_t)a + 1] - nd.cy);
        p2 += cq * (pos[3 * (int64_t)a + 2] - nd.cz);
    }
    mom[4 * (int64_t)id] = q;
    mom[4 * (int64_t)id + 1] = p0;
    mom[4 * (int64_t)id + 2] = p1;
    mom[4 * (int64_t)id + 3] = p2;
}

// compute_moments_rec, internal nodes of one depth (children done before)
__global__ void __launch_bounds__(kNT) k_tc_node_moments(const TNode *__restrict__ nodes,
                                                         const int *__restrict__ ids, int n,
                                                         double *__restrict__ mom) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int id = ids[t];
    const TNode nd = nodes[id];
    double q = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int c = id + 1; c < nd.skip; c = nodes[c].skip) {
        const TNode cn = nodes[c];
        const double cq = mom[4 * (int64_t)c];
        q += cq;
        p0 += mom[4 * (int64_t)c + 1] + cq * (cn.cx - nd.cx);
        p1 += mom[4 * (int64_t)c + 2] + cq * (cn.cy - nd.cy);
        p2 += mom[4 * (int64_t)c + 3] + cq * (cn.cz - nd.cz);
    }
    mom[4 * (int64_t)id] = q;
    mom[4 * (int64_t)id + 1] = p0;
    mom[4 * (int64_t)id + 2] = p1;
    mom[4 * (int64_t)id + 3] = p2;
}

struct BoxGeom {
    int nx, ny, nz;
    double dx, dy, dz, cx, cy, cz;
};

__device__ __forceinline__ void cell_pos(const BoxGeom &g, int i, int j, int k, double &x, double &y, double &z) {
    x = ((double)i + 0.5 - g.cx) * g.dx;
    y = ((double)j + 0.5 - g.cy) * g.dy;
    z = ((double)k + 0.5 - g.cz) * g.dz;
}

// eval_phi (mg_treecode.rs:192-259) at every boundary cell; writes bc and
// stamps phi.  The boundary cells come in 8 x 8 patches of one face per wave
// (targets, -1 = idle lane), so a wave's cells are close together.
//
// The 64 lanes of a wave hold 64 neighbouring boundary cells and walk the
// pre-order node array together: node n is visited by the lanes whose own
// traversal reaches it (lane.next == n).  A lane that opens n continues at
// n + 1, a lane that accepts n (or sums the leaf n) continues at skip[n].
// Lanes waiting further on accepted an ancestor, whose skip is >= skip[n], so
// the wave's next node is n + 1 if any lane opened n and skip[n] otherwise.
// Node, moment and charge loads are wave-uniform.  Each lane sees exactly its
// own traversal in its own order, and keeps one partial sum per open depth,
// so its value is the reference's nested recursive sum bit for bit.  The
// partial sums live in LDS, one row of kNT lanes per depth (dynamic LDS of
// (deepest node depth + 1) rows): indexed by the run-time depth they would
// otherwise go to per-lane scratch memory.
__global__ void __launch_bounds__(kNT) k_tc_eval(const TNode *__restrict__ nodes, int nnodes,
                                                 const double *__restrict__ mom, const double *__restrict__ pos,
                                                 const double *__restrict__ qp, BoxGeom g, double theta,
                                                 const int *__restrict__ targets, int ntargets,
                                                 double *__restrict__ bc, double *__restrict__ phi) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int cell = t < ntargets ? targets[t] : -1;
    const bool valid = cell >= 0;
    const int i = valid ? cell % g.nx : 0, j = valid ? (cell / g.nx) % g.ny : 0,
              k = valid ? cell / (g.nx * g.ny) : 0;
    double tx, ty, tz;
    cell_pos(g, i, j, k, tx, ty, tz);
    extern __shared__ double part_lds[];
    double *part = part_lds + threadIdx.x;   // part[d * kNT]: depth d of this lane
    double total = 0.0;
    bool done = !valid;
    int next = 0;       // this lane's next node
    int n = 0;          // the wave's node (uniform)
    while (n < nnodes) {
        n = __builtin_amdgcn_readfirstlane(n);
        const TNode nd = nodes[n];
        bool opened = false;
        if (!done && next == n) {
            const int d = nd.depth;
            double v = 0.0;
            bool accept;
            if (nd.count > 0) {
                accept = true;
                for (int a = nd.start; a < nd.start + nd.count; ++a) {
                    const double ex = tx - pos[3 * (int64_t)a], ey = ty - pos[3 * (int64_t)a + 1],
                                 ez = tz - pos[3 * (int64_t)a + 2];
                    const double rr2 = ex * ex + ey * ey + ez * ez;
                    if (rr2 > 0.0) v += qp[a] / sqrt(rr2);
                }
            } else {
                const double rx = tx - nd.cx, ry = ty - nd.cy, rz = tz - nd.cz;
                const double ddx = fmax(fabs(rx) - nd.half, 0.0), ddy = fmax(fabs(ry) - nd.half, 0.0),
                             ddz = fmax(fabs(rz) - nd.half, 0.0);
                const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                const double size = nd.half * 2.0;
                // the reference's test is size / sqrt(d2) < theta (sqrt(d2) > 0
                // iff d2 > 0); outside a relative band of 1e-12 around equality
                // the squared comparison decides it the same way (its rounding
                // is below 1e-15), inside the band the exact test runs
                if (d2 > 0.0) {
                    const double lhs = size * size, rhs2 = theta * theta * d2;
                    if (lhs < rhs2 * (1.0 - 1e-12)) accept = true;
                    else if (lhs > rhs2 * (1.0 + 1e-12)) accept = false;
                    else accept = size / sqrt(d2) < theta;
                } else {
                    accept = false;
                }
                if (accept) {
                    const double r = sqrt(rx * rx + ry * ry + rz * rz);
                    const double inv_r = 1.0 / r;
                    const double inv_r3 = inv_r * inv_r * inv_r;
                    const double pr = mom[4 * (int64_t)n + 1] * rx + mom[4 * (int64_t)n + 2] * ry +
                                      mom[4 * (int64_t)n + 3] * rz;
                    v = mom[4 * (int64_t)n] * inv_r + pr * inv_r3;
                }
            }
            if (!accept) {
                part[d * kNT] = 0.0;
                next = n + 1;
                opened = true;
            } else if (d == 0) {
                total = v;
                done = true;
            } else {
                part[(d - 1) * kNT] += v;
                for (int l = d - 1; l >= nd.skip_depth; --l) {
                    if (l == 0) {
                        total = part[0];
                        done = true;
                        break;
                    }
                    part[(l - 1) * kNT] += part[l * kNT];
                }
                next = nd.skip;
            }
        }
        n = __ballot(opened) ? n + 1 : nd.skip;
    }
    if (!valid) return;
    const int64_t id = idx3(i, j, k, g.nx, g.ny);
    const double val = -total / (4.0 * kPi);
    bc[id] = val;
    phi[id] = val;
}

// evaluate_dipole_bc (mg_treecode.rs:371-448): moments per block, then one
// thread sums the partials in block order.
__global__ void __launch_bounds__(kNT) k_dip_partial(const double *__restrict__ rhs, BoxGeom g, int k_lo, int nk,
                                                     double dvol, double *__restrict__ partial) {
    __shared__ double red[4][kNT];
    const int64_t plane = (int64_t)(g.nx - 2) * (g.ny - 2);
    const int64_t n = plane * nk;
    double q = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int k = k_lo + (int)(t / plane);
        const int64_t r = t % plane;
        const int j = 1 + (int)(r / (g.nx - 2));
        const int i = 1 + (int)(r % (g.nx - 2));
        double x, y, z;
        cell_pos(g, i, j, k, x, y, z);
        const double w = rhs[idx3(i, j, k, g.nx, g.ny)] * dvol;
        q += w;
        px += w * x;
        py += w * y;
        pz += w * z;
    }
    red[0][threadIdx.x] = q;
    red[1][threadIdx.x] = px;
    red[2][threadIdx.x] = py;
    red[3][threadIdx.x] = pz;
    __syncthreads();
    for (int s = kNT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_dip_final(const double *__restrict__ partial, int nb, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < 4; ++c) s[c] += partial[4 * b + c];
    for (int c = 0; c < 4; ++c) out[c] = s[c];
}

__global__ void __launch_bounds__(kNT) k_dip_eval(const double *__restrict__ mom, BoxGeom g,
                                                  double *__restrict__ bc, double *__restrict__ phi) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= boundary_count(g.nx, g.ny, g.nz)) return;
    int i, j, k;
    boundary_cell(t, g.nx, g.ny, g.nz, i, j, k);
    double x, y, z;
    cell_pos(g, i, j, k, x, y, z);
    const double q = mom[0], pxm = mom[1], pym = mom[2], pzm = mom[3];
    const double inv4pi = 1.0 / (4.0 * kPi);
    const double r2 = x * x + y * y + z * z;
    double v = 0.0;
    if (r2 > 0.0) {
        const double r = sqrt(r2);
        const double pr = pxm * x + pym * y + pzm * z;
        v = -inv4pi * (q / r + pr / (r * r * r));
    }
    const int64_t id = idx3(i, j, k, g.nx, g.ny);
    bc[id] = v;
    phi[id] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host: configuration
// ---------------------------------------------------------------------------
void mg_config_default(llgmi_mg_config &c) {
    std::memset(&c, 0, sizeof(c));
    c.pad_xy = 6;
    c.n_vac_z = 16;
    c.v_cycles = 4;
    c.v_cycles_max = 64;
    c.has_tol_abs = 0;
    c.tol_abs = 0.0;
    c.has_tol_rel = 1;
    c.tol_rel = 1e-6;
    c.pre_smooth = 2;
    c.post_smooth = 2;
    c.smoother = 0;
    c.omega = 2.0 / 3.0;
    c.sor_omega = 1.0;
    c.warm_start = 1;
    c.bc = 2;
    c.tree_theta = 0.6;
    c.tree_leaf = 64;
    c.tree_max_depth = 20;
    c.stencil = 2;
    c.prolong = 1;
    c.coarse_op = 1;
    c.iso27_alpha = 0.0;
}

namespace {
std::string trim(const std::string &s) {
    const auto b = s.find_first_not_of(" \t\r\n\v\f");
    if (b == std::string::npos) return std::string();
    const auto e = s.find_last_not_of(" \t\r\n\v\f");
    return s.substr(b, e - b + 1);
}
std::string lower(std::string s) {
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
bool env_str(const char *name, std::string &out) {
    const char *v = std::getenv(name);
    if (!v) return false;
    out = v;
    return true;
}
// str::trim().parse::<usize>()
bool env_usize(const char *name, int64_t &out) {
    std::string s;
    if (!env_str(name, s)) return false;
    s = trim(s);
    size_t p = 0;
    if (!s.empty() && s[0] == '+') p = 1;
    if (p >= s.size()) return false;
    uint64_t v = 0;
    for (; p < s.size(); ++p) {
        if (s[p] < '0' || s[p] > '9') return false;
        if (v > (UINT64_MAX - 9) / 10) return false;
        v = v * 10 + (uint64_t)(s[p] - '0');
    }
    if (v > (uint64_t)INT64_MAX) return false;
    out = (int64_t)v;
    return true;
}
// str::trim().parse::<f64>()
bool env_f64(const char *name, double &out) {
    std::string s;
    if (!env_str(name, s)) return false;
    s = trim(s);
    if (s.empty()) return false;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    out = v;
    return true;
}
double clampf(double v, double lo, double hi) {  // f64::clamp (NaN stays NaN)
    if (std::isnan(v)) return v;
    return v < lo ? lo : (v > hi ? hi : v);
}
}  // namespace

// HybridConfig::default (:283-293) and the boundary-scaling defaults (:2743-2760)
void mg_hybrid_config_default(llgmi_mg_hybrid_config &h) {
    std::memset(&h, 0, sizeof h);
    h.enabled = 0;
    h.radius_xy = 12;
    h.delta_v_cycles = 60;
    h.sigma_cells = 1.5;
    h.bscale = 0.0;
    h.bmargin = 2;
}

// HybridConfig::from_env (:296-356); LLG_DEMAG_MG_HYBRID_BSCALE / _BMARGIN are
// parsed once per process without trimming, as the reference's OnceLock does.
void mg_hybrid_config_from_env(llgmi_mg_hybrid_config &h) {
    mg_hybrid_config_default(h);
    std::string s;
    if (env_str("LLG_DEMAG_MG_HYBRID_ENABLE", s)) {
        const std::string t = lower(trim(s));
        h.enabled = (t == "1" || t == "true" || t == "yes" || t == "on") ? 1 : 0;
    }
    if (h.enabled) {
        int64_t u;
        double f;
        if (env_usize("LLG_DEMAG_MG_HYBRID_RADIUS", u)) h.radius_xy = u;
        if (env_usize("LLG_DEMAG_MG_HYBRID_DELTA_VCYCLES", u)) h.delta_v_cycles = std::max<int64_t>(u, 1);
        if (env_f64("LLG_DEMAG_MG_HYBRID_SIGMA", f)) {
            f = std::isnan(f) ? 0.0 : std::max(f, 0.0);   // f64::max ignores NaN
            h.sigma_cells = std::min(f, 32.0);
        }
    }
    static const double bscale = [] {
        const char *v = std::getenv("LLG_DEMAG_MG_HYBRID_BSCALE");
        if (!v || !*v) return 0.0;
        char *end = nullptr;
        const double d = std::strtod(v, &end);
        return (end && *end == '\0') ? d : 0.0;
    }();
    static const int64_t bmargin = [] {
        const char *v = std::getenv("LLG_DEMAG_MG_HYBRID_BMARGIN");
        if (!v || !*v) return (int64_t)2;
        const char *p = v + (*v == '+' ? 1 : 0);
        if (!*p) return (int64_t)2;
        uint64_t x = 0;
        for (; *p; ++p) {
            if (*p < '0' || *p > '9' || x > (UINT64_MAX - 9) / 10) return (int64_t)2;
            x = x * 10 + (uint64_t)(*p - '0');
        }
        return x > (uint64_t)INT64_MAX ? (int64_t)2 : (int64_t)x;
    }();
    h.bscale = bscale;
    h.bmargin = bmargin;
}

// DemagPoissonMGConfig::from_env (:188-267) + MGOperatorSettings::from_env (:741-794)
void mg_config_from_env(llgmi_mg_config &c) {
    mg_config_default(c);
    int64_t u;
    double f;
    std::string s;
    if (env_usize("LLG_DEMAG_MG_PAD_XY", u) || env_usize("LLG_DEMAG_MG_PAD_FACTOR_XY", u)) c.pad_xy = std::max<int64_t>(u, 1);
    if (env_usize("LLG_DEMAG_MG_NVAC_Z", u)) c.n_vac_z = std::max<int64_t>(u, 1);
    if (env_usize("LLG_DEMAG_MG_VCYCLES", u)) c.v_cycles = std::max<int64_t>(u, 1);
    if (env_usize("LLG_DEMAG_MG_VCYCLES_MAX", u)) c.v_cycles_max = std::max<int64_t>(u, 1);
    if (env_f64("LLG_DEMAG_MG_TOL_ABS", f)) { c.has_tol_abs = 1; c.tol_abs = std::fmax(f, 0.0); }
    if (env_f64("LLG_DEMAG_MG_TOL_REL", f)) { c.has_tol_rel = 1; c.tol_rel = std::fmin(std::fmax(f, 0.0), 1.0); }
    if (env_usize("LLG_DEMAG_MG_PRE_SMOOTH", u)) c.pre_smooth = std::max<int64_t>(u, 1);
    if (env_usize("LLG_DEMAG_MG_POST_SMOOTH", u)) c.post_smooth = std::max<int64_t>(u, 1);
    if (env_f64("LLG_DEMAG_MG_OMEGA", f)) c.omega = clampf(f, 0.05, 1.0);
    if (env_f64("LLG_DEMAG_MG_SOR_OMEGA", f)) c.sor_omega = clampf(f, 0.2, 1.95);
    if (env_str("LLG_DEMAG_MG_SMOOTHER", s)) {   // MGSmoother::from_str (:97-112)
        const std::string t = lower(trim(s));
        if (t == "jacobi" || t == "wj" || t == "weighted_jacobi") c.smoother = 0;
        else if (t == "rbgs" || t == "redblack" || t == "red_black" || t == "red_black_sor" || t == "sor") c.smoother = 1;
    }
    if (env_str("LLG_DEMAG_MG_WARM_START", s))
        c.warm_start = (s == "1" || s == "true" || s == "yes" || s == "on") ? 1 : 0;
    if (env_str("LLG_DEMAG_MG_BC", s)) {
        const std::string t = lower(trim(s));
        if (t == "0" || t == "zero" || t == "dirichlet0" || t == "dirichlet_zero") c.bc = 0;
        else if (t == "dipole" || t == "dirichlet_dipole" || t == "dirichletdipole") c.bc = 1;
        else if (t == "tree" || t == "treecode" || t == "bh" || t == "fmm") c.bc = 2;
    }
    if (env_f64("LLG_DEMAG_MG_TREE_THETA", f)) c.tree_theta = clampf(f, 0.2, 1.5);
    if (env_usize("LLG_DEMAG_MG_TREE_LEAF", u)) c.tree_leaf = std::min<int64_t>(std::max<int64_t>(u, 8), 4096);
    if (env_usize("LLG_DEMAG_MG_TREE_MAX_DEPTH", u)) c.tree_max_depth = std::min<int64_t>(std::max<int64_t>(u, 4), 64);
    if (env_str("LLG_DEMAG_MG_STENCIL", s)) {
        const std::string t = lower(trim(s));
        if (t == "7" || t == "7pt" || t == "seven" || t == "sevenpoint") c.stencil = 0;
        else if (t == "iso9" || t == "9" || t == "9pt" || t == "mehrstellen9") c.stencil = 1;
        else c.stencil = 2;
    }
    if (env_str("LLG_DEMAG_MG_PROLONG", s)) {
        const std::string t = lower(trim(s));
        c.prolong = (t == "inject" || t == "injection" || t == "pc" || t == "piecewiseconstant") ? 0 : 1;
    }
    if (env_str("LLG_DEMAG_MG_COARSE_OP", s)) {
        const std::string t = lower(trim(s));
        c.coarse_op = (t == "rediscretize" || t == "re" || t == "rd") ? 0 : 1;
    }
    if (env_f64("LLG_DEMAG_MG_ISO27_FLUX_ALPHA", f) && std::isfinite(f) && f > 0.0) c.iso27_alpha = f;
}

// rb_allowed_for_op / sanitize_cfg_for_op (:1400-1417)
void mg_sanitize(llgmi_mg_config &c) {
    const bool rb_ok = c.stencil == 0 && c.coarse_op == 0 && c.prolong == 0;
    if (c.smoother == 1 && !rb_ok) c.smoother = 0;
}

void mg_validate(const llgmi_mg_config &c) {
    auto bad = [](const std::string &m) { throw Error(1, "llgmi_mg_config: " + m); };
    if (c.pad_xy < 0 || c.n_vac_z < 0 || c.v_cycles < 0 || c.v_cycles_max < 0 || c.pre_smooth < 0 ||
        c.post_smooth < 0)
        bad("counts must be >= 0");
    if (c.smoother != 0 && c.smoother != 1) bad("smoother must be 0 or 1");
    if (c.bc < 0 || c.bc > 2) bad("bc must be 0, 1 or 2");
    if (c.stencil < 0 || c.stencil > 2) bad("stencil must be 0, 1 or 2");
    if (c.prolong != 0 && c.prolong != 1) bad("prolong must be 0 or 1");
    if (c.coarse_op != 0 && c.coarse_op != 1) bad("coarse_op must be 0 or 1");
    if (c.tree_leaf < 1) bad("tree_leaf must be >= 1");
    if (c.tree_max_depth < 0 || c.tree_max_depth > kMaxDepth) bad("tree_max_depth must be in [0, 64]");
    if (!std::isfinite(c.omega) || !std::isfinite(c.sor_omega) || !std::isfinite(c.tree_theta))
        bad("omega, sor_omega and tree_theta must be finite");
    if (c.pad_xy > (1 << 20) || c.n_vac_z > (1 << 20)) bad("padding too large");
}

// ---------------------------------------------------------------------------
// Host: stencils (demag_poisson_mg.rs:840-1189)
// ---------------------------------------------------------------------------
namespace {
void push(MGStencil &s, int a, int b, int c, double w) {
    s.offs.push_back({a, b, c});
    s.coeffs.push_back(w);
}

MGStencil seven_point(double dx, double dy, double dz) {
    const double sx = 1.0 / (dx * dx), sy = 1.0 / (dy * dy), sz = 1.0 / (dz * dz);
    MGStencil s;
    s.center = -2.0 * (sx + sy + sz);
    push(s, 1, 0, 0, sx);
    push(s, -1, 0, 0, sx);
    push(s, 0, 1, 0, sy);
    push(s, 0, -1, 0, sy);
    push(s, 0, 0, 1, sz);
    push(s, 0, 0, -1, sz);
    s.diag = -s.center;
    return s;
}

MGStencil iso9_plus_z(double dx, double dy, double dz) {
    const double rel = std::fabs(dx - dy) / std::fmax(std::fmax(dx, dy), 1e-30);
    if (rel > 1e-6) return seven_point(dx, dy, dz);
    const double h = 0.5 * (dx + dy);
    const double inv_h2 = 1.0 / (h * h);
    const double cz = 1.0 / (dz * dz);
    const double c_axis = (2.0 / 3.0) * inv_h2;
    const double c_diag = (1.0 / 6.0) * inv_h2;
    MGStencil s;
    s.center = -(10.0 / 3.0) * inv_h2 - 2.0 * cz;
    push(s, 1, 0, 0, c_axis);
    push(s, -1, 0, 0, c_axis);
    push(s, 0, 1, 0, c_axis);
    push(s, 0, -1, 0, c_axis);
    push(s, 1, 1, 0, c_diag);
    push(s, 1, -1, 0, c_diag);
    push(s, -1, 1, 0, c_diag);
    push(s, -1, -1, 0, c_diag);
    push(s, 0, 0, 1, cz);
    push(s, 0, 0, -1, cz);
    s.diag = -s.center;
    return s;
}

MGStencil iso27(double dx, double dy, double dz, double alpha_in) {
    double sx1 = 0.0, sy1 = 0.0, sz1 = 0.0;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                if (std::abs(di) + std::abs(dj) + std::abs(dk) == 1) continue;
                const double ddx = (double)di * dx, ddy = (double)dj * dy, ddz = (double)dk * dz;
                const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= 0.0) continue;
                const double w1 = 1.0 / d2;
                sx1 += w1 * ddx * ddx;
                sy1 += w1 * ddy * ddy;
                sz1 += w1 * ddz * ddz;
            }
    const bool have = std::isfinite(alpha_in) && alpha_in > 0.0;
    double amax = INFINITY;
    if (sx1 > 0.0) amax = std::fmin(amax, 2.0 / sx1);
    if (sy1 > 0.0) amax = std::fmin(amax, 2.0 / sy1);
    if (sz1 > 0.0) amax = std::fmin(amax, 2.0 / sz1);
    if (std::isfinite(amax)) amax *= 0.999;
    const double adef = std::isfinite(amax) ? std::fmax(0.99 * amax, 0.0) : 0.0;
    double alpha = have ? alpha_in : adef;
    if (std::isfinite(amax) && alpha > amax) alpha = std::fmax(amax, 0.0);
    const double wfx = (2.0 - alpha * sx1) / (2.0 * dx * dx);
    const double wfy = (2.0 - alpha * sy1) / (2.0 * dy * dy);
    const double wfz = (2.0 - alpha * sz1) / (2.0 * dz * dz);
    MGStencil s;
    push(s, 1, 0, 0, wfx);
    push(s, -1, 0, 0, wfx);
    push(s, 0, 1, 0, wfy);
    push(s, 0, -1, 0, wfy);
    push(s, 0, 0, 1, wfz);
    push(s, 0, 0, -1, wfz);
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                if (std::abs(di) + std::abs(dj) + std::abs(dk) <= 1) continue;
                const double ddx = (double)di * dx, ddy = (double)dj * dy, ddz = (double)dk * dz;
                const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= 0.0) continue;
                push(s, di, dj, dk, alpha / d2);
            }
    double sum = 0.0;
    for (double c : s.coeffs) sum += c;
    s.center = -sum;
    s.diag = sum;
    return s;
}

inline int64_t hidx(int64_t i, int64_t j, int64_t k, int64_t nx, int64_t ny) { return (k * ny + j) * nx + i; }

double apply_at(const MGStencil &s, const std::vector<double> &phi, int nx, int ny, int nz, int i, int j,
                int k) {
    double sum = 0.0;
    for (size_t q = 0; q < s.coeffs.size(); ++q) {
        const int ii = std::min(std::max(i + s.offs[q][0], 0), nx - 1);
        const int jj = std::min(std::max(j + s.offs[q][1], 0), ny - 1);
        const int kk = std::min(std::max(k + s.offs[q][2], 0), nz - 1);
        sum += s.coeffs[q] * phi[hidx(ii, jj, kk, nx, ny)];
    }
    return s.center * phi[hidx(i, j, k, nx, ny)] + sum;
}

void interp_1d_h(int ic, int ri, int nc, int r, int &i0o, int &i1o, double &w0, double &w1) {
    const int i0 = std::min(ic, nc - 1);
    if (r == 1) { i0o = i1o = i0; w0 = 1.0; w1 = 0.0; return; }
    const int i1 = ri == 0 ? (i0 > 0 ? i0 - 1 : 0) : std::min(i0 + 1, nc - 1);
    if (i1 == i0) { i0o = i1o = i0; w0 = 1.0; w1 = 0.0; }
    else { i0o = i0; i1o = i1; w0 = 0.75; w1 = 0.25; }
}
}  // namespace

MGStencil mg_stencil_from_kind(int kind, double dx, double dy, double dz, double alpha) {
    if (kind == 0) return seven_point(dx, dy, dz);
    if (kind == 1) return iso9_plus_z(dx, dy, dz);
    return iso27(dx, dy, dz, alpha);
}

// Stencil3D::galerkin_coarsen (:1087-1189).  Only the fine cells of the probe
// box that restrict into the centre coarse cell are evaluated; each value is
// computed exactly as in the full-box evaluation.
MGStencil mg_galerkin(const MGStencil &fine, int rx, int ry, int rz, bool trilinear) {
    const int nc = 9, c0 = nc / 2;
    const int nfx = nc * rx, nfy = nc * ry, nfz = nc * rz;
    std::vector<double> phic((size_t)nc * nc * nc), phif((size_t)nfx * nfy * nfz);
    std::map<std::tuple<int, int, int>, double> map;
    const double norm = 1.0 / (double)(rx * ry * rz);
    for (int kz = 0; kz < nc; ++kz)
        for (int jy = 0; jy < nc; ++jy)
            for (int ix = 0; ix < nc; ++ix) {
                std::fill(phic.begin(), phic.end(), 0.0);
                phic[hidx(ix, jy, kz, nc, nc)] = 1.0;
                std::fill(phif.begin(), phif.end(), 0.0);
                if (!trilinear) {
                    for (int fk = 0; fk < rz; ++fk)
                        for (int fj = 0; fj < ry; ++fj)
                            for (int fi = 0; fi < rx; ++fi)
                                phif[hidx(ix * rx + fi, jy * ry + fj, kz * rz + fk, nfx, nfy)] = 1.0;
                } else {
                    for (int k = 0; k < nfz; ++k) {
                        int k0, k1; double wk0, wk1;
                        interp_1d_h(k / rz, k % rz, nc, rz, k0, k1, wk0, wk1);
                        for (int j = 0; j < nfy; ++j) {
                            int j0, j1; double wj0, wj1;
                            interp_1d_h(j / ry, j % ry, nc, ry, j0, j1, wj0, wj1);
                            for (int i = 0; i < nfx; ++i) {
                                int i0, i1; double wi0, wi1;
                                interp_1d_h(i / rx, i % rx, nc, rx, i0, i1, wi0, wi1);
                                auto c = [&](int a, int b, int d) { return phic[hidx(a, b, d, nc, nc)]; };
                                double v = 0.0;
                                v += wi0 * wj0 * wk0 * c(i0, j0, k0);
                                v += wi1 * wj0 * wk0 * c(i1, j0, k0);
                                v += wi0 * wj1 * wk0 * c(i0, j1, k0);
                                v += wi1 * wj1 * wk0 * c(i1, j1, k0);
                                v += wi0 * wj0 * wk1 * c(i0, j0, k1);
                                v += wi1 * wj0 * wk1 * c(i1, j0, k1);
                                v += wi0 * wj1 * wk1 * c(i0, j1, k1);
                                v += wi1 * wj1 * wk1 * c(i1, j1, k1);
                                phif[hidx(i, j, k, nfx, nfy)] = v;
                            }
                        }
                    }
                }
                double sum = 0.0;
                for (int fk = 0; fk < rz; ++fk)
                    for (int fj = 0; fj < ry; ++fj)
                        for (int fi = 0; fi < rx; ++fi)
                            sum += apply_at(fine, phif, nfx, nfy, nfz, c0 * rx + fi, c0 * ry + fj, c0 * rz + fk);
                const double coeff = norm * sum;
                if (std::fabs(coeff) > 1e-14) map[std::make_tuple(ix - c0, jy - c0, kz - c0)] = coeff;
            }
    MGStencil out;
    for (const auto &kv : map) {
        const auto &key = kv.first;
        if (key == std::make_tuple(0, 0, 0)) out.center = kv.second;
        else push(out, std::get<0>(key), std::get<1>(key), std::get<2>(key), kv.second);
    }
    out.diag = -out.center;
    double abs_sum = 0.0;
    size_t neg = 0;
    for (double c : out.coeffs) {
        abs_sum += std::fabs(c);
        if (c < 0.0) ++neg;
    }
    const double rho = out.diag > 0.0 ? abs_sum / out.diag : INFINITY;
    if (out.diag <= 0.0 || rho > 2.0 || neg > out.coeffs.size() / 2) return MGStencil{};
    return out;
}

// ---------------------------------------------------------------------------
// Host: treecode octree (BarnesHutTree, mg_treecode.rs:71-156)
// ---------------------------------------------------------------------------
// The structure that sequential insertion produces depends only on the charge
// set: a node below max_depth is split iff it holds more than leaf_size
// charges, children exist for non-empty octants, and a leaf lists its charges
// in increasing charge order.  The build below partitions stably level by
// level, which yields exactly that tree, and lays it out in pre-order with
// children in octant order (the evaluation order of eval_node).
struct MGSolver::Tree {
    int64_t n = 0;
    int leaf = -1, max_depth = -1;
    bool valid = false;
    int nnodes = 0;
    int deepest = 0;   // largest node depth of the built tree
    // compaction
    int *d_counts = nullptr, *d_offsets = nullptr, *d_meta = nullptr, *h_meta = nullptr;
    int64_t cap_blocks = 0;
    int *d_idx = nullptr;
    double *d_q = nullptr;
    int64_t cap_cand = 0;
    int *d_cached = nullptr;
    int64_t cap_cached = 0;
    int *d_targets = nullptr;   // boundary cells in face patches (-1 = idle lane)
    int ntargets = 0;
    // tree
    TNode *d_nodes = nullptr;
    double *d_mom = nullptr, *d_pos = nullptr, *d_qp = nullptr;
    int *d_ord = nullptr, *d_leaves = nullptr, *d_internal = nullptr;
    int nleaves = 0;
    std::vector<std::pair<int, int>> internal_by_depth;  // (offset, count), depth order

    void free_tree() {
        for (void *p : {(void *)d_nodes, (void *)d_mom, (void *)d_pos, (void *)d_qp, (void *)d_ord,
                        (void *)d_leaves, (void *)d_internal})
            if (p) (void)hipFree(p);
        d_nodes = nullptr; d_mom = d_pos = d_qp = nullptr; d_ord = d_leaves = d_internal = nullptr;
        valid = false;
    }
    ~Tree() {
        free_tree();
        for (void *p : {(void *)d_counts, (void *)d_offsets, (void *)d_meta, (void *)d_idx, (void *)d_q,
                        (void *)d_cached, (void *)d_targets})
            if (p) (void)hipFree(p);
        if (h_meta) (void)hipHostFree(h_meta);
    }
};

namespace {
struct HostTreeBuilder {
    const std::vector<double> &pos;   // 3 per charge (raster rank)
    int leaf, max_depth;
    std::vector<int> ord, tmp;
    std::vector<TNode> nodes;
    std::vector<int> leaves;
    std::vector<std::vector<int>> internal;   // by depth

    HostTreeBuilder(const std::vector<double> &p, int l, int md) : pos(p), leaf(l), max_depth(md) {}

    static int octant(const double c[3], const double *p) {
        int o = 0;
        if (p[0] >= c[0]) o |= 1;
        if (p[1] >= c[1]) o |= 2;
        if (p[2] >= c[2]) o |= 4;
        return o;
    }

    void emit(int lo, int hi, const double c[3], double half, int depth) {
        const int id = (int)nodes.size();
        nodes.push_back(TNode{c[0], c[1], c[2], half, -1, depth, 0, 0, 0, 0});
        if (depth >= max_depth || hi - lo <= leaf) {
            nodes[id].start = lo;
            nodes[id].count = hi - lo;
            leaves.push_back(id);
        } else {
            int cnt[8] = {0}, at[8];
            for (int r = lo; r < hi; ++r) cnt[octant(c, &pos[3 * (size_t)ord[r]])]++;
            at[0] = lo;
            for (int o = 1; o < 8; ++o) at[o] = at[o - 1] + cnt[o - 1];
            for (int r = lo; r < hi; ++r) tmp[at[octant(c, &pos[3 * (size_t)ord[r]])]++] = ord[r];
            std::copy(tmp.begin() + lo, tmp.begin() + hi, ord.begin() + lo);
            if ((int)internal.size() <= depth) internal.resize(depth + 1);
            internal[depth].push_back(id);
            const double ch = half * 0.5;
            int b = lo;
            for (int o = 0; o < 8; ++o) {
                if (cnt[o] == 0) continue;
                const double cc[3] = {c[0] + ((o & 1) ? ch : -ch), c[1] + ((o & 2) ? ch : -ch),
                                      c[2] + ((o & 4) ? ch : -ch)};
                emit(b, b + cnt[o], cc, ch, depth + 1);
                b += cnt[o];
            }
        }
        nodes[id].skip = (int)nodes.size();
    }
};

template <typename P>
void dev_alloc(P *&p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    LLGMI_HIP(hipMalloc((void **)&p, sizeof(P) * std::max<size_t>(count, 1)));
}
}  // namespace

// ---------------------------------------------------------------------------
// Host: solver
// ---------------------------------------------------------------------------
MGSolver::MGSolver(const llgmi_grid &g, const llgmi_mg_config &cfg_in, hipStream_t stream, Profiler *prof)
    : g_(g), cfg_(cfg_in), stream_(stream), prof_(prof) {
    mg_validate(cfg_);
    mg_sanitize(cfg_);
    if (const char *e = std::getenv("LLGMI_MG_GRAPH")) use_graphs_ = std::string(e) != "0";
    if (const char *e = std::getenv("LLGMI_MG_RB_FUSED")) rb_fused_ = std::string(e) != "0";
    if (const char *e = std::getenv("LLGMI_MG_RBZ")) rbz_ = std::max(1, std::atoi(e));
    try {
        // new_with_operator (:1425-1577)
        const int64_t nx = std::max<int64_t>(g.nx, 1), ny = std::max<int64_t>(g.ny, 1);
        const int64_t pad = std::max<int64_t>(cfg_.pad_xy, 1);
        int64_t px = nx + 2 * pad, py = ny + 2 * pad;
        if (px % 2 == 1) ++px;
        if (py % 2 == 1) ++py;
        const int64_t nvac = std::max<int64_t>(cfg_.n_vac_z, 1);
        int64_t pz = 1 + 2 * nvac;
        if (pz % 2 == 1) ++pz;
        if (pz % 4 != 0) pz = ((pz + 3) / 4) * 4;
        if (px * py * pz >= ((int64_t)1 << 31)) throw Error(3, "multigrid box beyond 2^31 cells");
        px_ = (int)px; py_ = (int)py; pz_ = (int)pz;
        offx_ = (int)((px - nx) / 2);
        offy_ = (int)((py - ny) / 2);
        offz_ = (int)nvac;
        int lx = px_, ly = py_, lz = pz_;
        double dx = g.dx, dy = g.dy, dz = g.dz;
        lev_.push_back(Level{lx, ly, lz, dx, dy, dz});
        for (;;) {
            const bool cxy = lx >= 8 && ly >= 8 && lx % 2 == 0 && ly % 2 == 0;
            const bool cz = lz >= 8 && lz % 2 == 0;
            if (!cxy && !cz) break;
            if (cxy) { lx /= 2; ly /= 2; dx *= 2.0; dy *= 2.0; }
            if (cz) { lz /= 2; dz *= 2.0; }
            lev_.push_back(Level{lx, ly, lz, dx, dy, dz});
            if (lev_.size() > 32) break;
        }
        lev_[0].st = mg_stencil_from_kind(cfg_.stencil, lev_[0].dx, lev_[0].dy, lev_[0].dz, cfg_.iso27_alpha);
        for (size_t l = 1; l < lev_.size(); ++l) {
            Level &F = lev_[l - 1], &C = lev_[l];
            if (cfg_.coarse_op == 0) {
                C.st = mg_stencil_from_kind(cfg_.stencil, C.dx, C.dy, C.dz, cfg_.iso27_alpha);
            } else {
                C.st = mg_galerkin(F.st, F.nx / C.nx, F.ny / C.ny, F.nz / C.nz, cfg_.prolong == 1);
                if (C.st.diag <= 0.0 || C.st.offs.empty())
                    C.st = mg_stencil_from_kind(cfg_.stencil, C.dx, C.dy, C.dz, cfg_.iso27_alpha);
            }
        }
        for (auto &L : lev_) {
            if (L.st.coeffs.size() > (size_t)kMaxSt) throw Error(3, "multigrid stencil too wide");
            const size_t n = (size_t)L.nx * L.ny * L.nz;
            LLGMI_HIP(hipMalloc((void **)&L.phi[0], sizeof(double) * n));
            LLGMI_HIP(hipMalloc((void **)&L.rhs, sizeof(double) * n));
            LLGMI_HIP(hipMemsetAsync(L.phi[0], 0, sizeof(double) * n, stream_));
            LLGMI_HIP(hipMemsetAsync(L.rhs, 0, sizeof(double) * n, stream_));
            static const int o7[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
            L.seven = L.st.coeffs.size() == 6;
            for (size_t q = 0; L.seven && q < 6; ++q)
                for (int a = 0; a < 3; ++a) L.seven = L.seven && L.st.offs[q][a] == o7[q][a];
            DevStencil h{};
            h.center = L.st.center;
            h.diag = L.st.diag;
            h.n = (int)L.st.coeffs.size();
            for (int q = 0; q < h.n; ++q) {
                h.c[q] = L.st.coeffs[q];
                for (int a = 0; a < 3; ++a) h.off[q][a] = L.st.offs[q][a];
            }
            LLGMI_HIP(hipMalloc(&L.d_st, sizeof(DevStencil)));
            LLGMI_HIP(hipMemcpy(L.d_st, &h, sizeof(DevStencil), hipMemcpyHostToDevice));
        }
        LLGMI_HIP(hipFuncSetAttribute((const void *)k_mg_rb_tb, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(2 * sizeof(double) * TB_MAXZ * TBW * TBW)));
        const size_t n0 = (size_t)px_ * py_ * pz_;
        LLGMI_HIP(hipMalloc((void **)&bc_, sizeof(double) * n0));
        LLGMI_HIP(hipMemsetAsync(bc_, 0, sizeof(double) * n0, stream_));
        LLGMI_HIP(hipMalloc((void **)&d_partial_, sizeof(double) * 4 * kRedBlocks));
        LLGMI_HIP(hipMalloc((void **)&d_scalars_, sizeof(double) * 8));
        LLGMI_HIP(hipHostMalloc((void **)&h_scalars_, sizeof(double) * 8, hipHostMallocDefault));
        tree_ = new Tree();
        LLGMI_HIP(hipHostMalloc((void **)&tree_->h_meta, sizeof(int) * 2, hipHostMallocDefault));
        LLGMI_HIP(hipMalloc((void **)&tree_->d_meta, sizeof(int) * 2));
        LLGMI_HIP(hipStreamSynchronize(stream_));
    } catch (...) {
        release();
        throw;
    }
    stats_.levels = (int32_t)lev_.size();
    stats_.box[0] = px_;
    stats_.box[1] = py_;
    stats_.box[2] = pz_;
    stats_.last_residual = -1.0;
}

MGSolver::~MGSolver() { release(); }

void MGSolver::release() {
    for (auto &kv : graphs_)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    graphs_.clear();
    for (auto &L : lev_) {
        for (double *p : {L.phi[0], L.phi[1], L.rhs})
            if (p) (void)hipFree(p);
        if (L.d_st) (void)hipFree(L.d_st);
        L.phi[0] = L.phi[1] = L.rhs = nullptr;
        L.d_st = nullptr;
    }
    lev_.clear();
    if (bc_) (void)hipFree(bc_);
    if (d_partial_) (void)hipFree(d_partial_);
    if (d_scalars_) (void)hipFree(d_scalars_);
    if (h_scalars_) (void)hipHostFree(h_scalars_);
    bc_ = d_partial_ = d_scalars_ = h_scalars_ = nullptr;
    if (d_screen_w_) (void)hipFree(d_screen_w_);
    d_screen_w_ = nullptr;
    screen_cap_ = 0;
    delete tree_;
    tree_ = nullptr;
}

// same_structure (:1585-1593): grid, padding and operator
bool MGSolver::same_structure(const llgmi_grid &g, const llgmi_mg_config &c) const {
    const bool alpha_same = cfg_.iso27_alpha == c.iso27_alpha || (!(cfg_.iso27_alpha > 0.0) && !(c.iso27_alpha > 0.0));
    return g_.nx == g.nx && g_.ny == g.ny && g_.dx == g.dx && g_.dy == g.dy && g_.dz == g.dz &&
           cfg_.pad_xy == c.pad_xy && cfg_.n_vac_z == c.n_vac_z && cfg_.stencil == c.stencil &&
           cfg_.prolong == c.prolong && cfg_.coarse_op == c.coarse_op && alpha_same;
}

// apply_cfg (:1579-1583): runtime knobs; the operator stays as built
void MGSolver::apply_cfg(const llgmi_mg_config &c) {
    mg_validate(c);
    llgmi_mg_config k = c;
    k.stencil = cfg_.stencil;
    k.prolong = cfg_.prolong;
    k.coarse_op = cfg_.coarse_op;
    k.iso27_alpha = cfg_.iso27_alpha;
    mg_sanitize(k);
    cfg_ = k;
}

void MGSolver::build_rhs(const void *m, bool f32, double ms) {
    RhsGeom rg{px_, py_, pz_, offx_, offy_, offz_, (int)g_.nx, (int)g_.ny, lev_[0].dx, lev_[0].dy, lev_[0].dz, ms};
    const int k_lo = std::max(1, offz_ - 1), k_hi = std::min(pz_ - 2, offz_ + 1);
    const int nk = k_hi - k_lo + 1;
    const int64_t n = (int64_t)(px_ - 2) * (py_ - 2) * nk;
    const int64_t nm = g_.nx * g_.ny;
    if (f32)
        hipLaunchKernelGGL(k_mg_rhs<float>, dim3(nblocks(n)), dim3(kNT), 0, stream_, (const float *)m, nm, rg,
                           k_lo, nk, lev_[0].rhs);
    else
        hipLaunchKernelGGL(k_mg_rhs<double>, dim3(nblocks(n)), dim3(kNT), 0, stream_, (const double *)m, nm, rg,
                           k_lo, nk, lev_[0].rhs);
    LLGMI_LAUNCH_CHECK();
}

void MGSolver::stamp(int l) {
    Level &L = lev_[l];
    hipLaunchKernelGGL(k_mg_stamp, dim3(nblocks(boundary_count(L.nx, L.ny, L.nz))), dim3(kNT), 0, stream_,
                       L.phi[L.cur], l == 0 ? bc_ : nullptr, L.nx, L.ny, L.nz);
    LLGMI_LAUNCH_CHECK();
}

void MGSolver::treecode_bc() {
    Tree &T = *tree_;
    Level &L = lev_[0];
    const int k_lo = std::max(1, offz_ - 1), k_hi = std::min(pz_ - 2, offz_ + 1);
    const int64_t ncand = (int64_t)(px_ - 2) * (py_ - 2) * (k_hi - k_lo + 1);
    const int64_t nb = nblocks(ncand);
    if (nb > T.cap_blocks) {
        dev_alloc(T.d_counts, (size_t)nb);
        dev_alloc(T.d_offsets, (size_t)nb);
        T.cap_blocks = nb;
    }
    if (ncand > T.cap_cand) {
        dev_alloc(T.d_idx, (size_t)ncand);
        dev_alloc(T.d_q, (size_t)ncand);
        T.cap_cand = ncand;
    }
    const double dvol = L.dx * L.dy * L.dz;
    hipLaunchKernelGGL(k_tc_count, dim3((unsigned)nb), dim3(kNT), 0, stream_, L.rhs, px_, py_, k_lo, ncand,
                       T.d_counts);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tc_scan, dim3(1), dim3(1024), 0, stream_, T.d_counts, (int)nb, T.d_offsets, T.d_meta);
    LLGMI_LAUNCH_CHECK();
    LLGMI_HIP(hipMemsetAsync(T.d_meta, 0, sizeof(int), stream_));
    hipLaunchKernelGGL(k_tc_write, dim3((unsigned)nb), dim3(kNT), 0, stream_, L.rhs, px_, py_, k_lo, ncand,
                       T.d_offsets, dvol, T.d_idx, T.d_q, T.d_cached, T.valid ? (int)T.n : 0, T.d_meta);
    LLGMI_LAUNCH_CHECK();
    LLGMI_HIP(hipMemcpyAsync(T.h_meta, T.d_meta, sizeof(int) * 2, hipMemcpyDeviceToHost, stream_));
    LLGMI_HIP(hipStreamSynchronize(stream_));
    const int64_t n = T.h_meta[1];
    stats_.n_charges = (int32_t)n;
    stats_.tree_rebuilt = 0;
    if (n == 0) {   // evaluate_treecode_bc returns with bc = 0
        LLGMI_HIP(hipMemsetAsync(bc_, 0, sizeof(double) * (size_t)px_ * py_ * pz_, stream_));
        stamp(0);
        return;
    }
    const bool rebuild = !T.valid || T.h_meta[0] != 0 || n != T.n || T.leaf != (int)cfg_.tree_leaf ||
                         T.max_depth != (int)cfg_.tree_max_depth;
    if (rebuild) {
        std::vector<int> idx((size_t)n);
        LLGMI_HIP(hipMemcpy(idx.data(), T.d_idx, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
        // charge positions, build_charges_from_rhs (mg_treecode.rs:280-296)
        const double cx = (double)px_ * 0.5, cy = (double)py_ * 0.5, cz = (double)pz_ * 0.5;
        std::vector<double> pos(3 * (size_t)n);
        for (int64_t a = 0; a < n; ++a) {
            const int64_t cell = idx[(size_t)a];
            const int64_t i = cell % px_, j = (cell / px_) % py_, k = cell / ((int64_t)px_ * py_);
            pos[3 * a] = ((double)i + 0.5 - cx) * L.dx;
            pos[3 * a + 1] = ((double)j + 0.5 - cy) * L.dy;
            pos[3 * a + 2] = ((double)k + 0.5 - cz) * L.dz;
        }
        const double lx = (double)px_ * L.dx, ly = (double)py_ * L.dy, lz = (double)pz_ * L.dz;
        const double root_half = 0.5 * std::fmax(std::fmax(lx, ly), lz) + 1e-12;
        HostTreeBuilder b(pos, (int)std::min<int64_t>(cfg_.tree_leaf, INT32_MAX), (int)cfg_.tree_max_depth);
        b.ord.resize((size_t)n);
        b.tmp.resize((size_t)n);
        for (int64_t a = 0; a < n; ++a) b.ord[(size_t)a] = (int)a;
        const double rc[3] = {0.0, 0.0, 0.0};
        b.emit(0, (int)n, rc, root_half, 0);
        for (auto &nd : b.nodes) nd.skip_depth = nd.skip < (int)b.nodes.size() ? b.nodes[(size_t)nd.skip].depth : 0;
        std::vector<double> pos_leaf(3 * (size_t)n);
        for (int64_t a = 0; a < n; ++a)
            for (int c = 0; c < 3; ++c) pos_leaf[3 * a + c] = pos[3 * (size_t)b.ord[(size_t)a] + c];
        std::vector<int> internal;
        T.internal_by_depth.clear();
        for (auto &lvl : b.internal) {
            T.internal_by_depth.push_back({(int)internal.size(), (int)lvl.size()});
            internal.insert(internal.end(), lvl.begin(), lvl.end());
        }
        T.free_tree();
        T.nnodes = (int)b.nodes.size();
        T.nleaves = (int)b.leaves.size();
        T.deepest = 0;
        for (const auto &nd : b.nodes) T.deepest = std::max(T.deepest, nd.depth);
        dev_alloc(T.d_nodes, b.nodes.size());
        dev_alloc(T.d_mom, 4 * b.nodes.size());
        dev_alloc(T.d_pos, pos_leaf.size());
        dev_alloc(T.d_qp, (size_t)n);
        dev_alloc(T.d_ord, (size_t)n);
        dev_alloc(T.d_leaves, b.leaves.size());
        dev_alloc(T.d_internal, internal.size());
        LLGMI_HIP(hipMemcpy(T.d_nodes, b.nodes.data(), sizeof(TNode) * b.nodes.size(), hipMemcpyHostToDevice));
        LLGMI_HIP(hipMemcpy(T.d_pos, pos_leaf.data(), sizeof(double) * pos_leaf.size(), hipMemcpyHostToDevice));
        LLGMI_HIP(hipMemcpy(T.d_ord, b.ord.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
        LLGMI_HIP(hipMemcpy(T.d_leaves, b.leaves.data(), sizeof(int) * b.leaves.size(), hipMemcpyHostToDevice));
        if (!internal.empty())
            LLGMI_HIP(hipMemcpy(T.d_internal, internal.data(), sizeof(int) * internal.size(), hipMemcpyHostToDevice));
        if (n > T.cap_cached) {
            dev_alloc(T.d_cached, (size_t)n);
            T.cap_cached = n;
        }
        LLGMI_HIP(hipMemcpyAsync(T.d_cached, T.d_idx, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, stream_));
        T.n = n;
        T.leaf = (int)cfg_.tree_leaf;
        T.max_depth = (int)cfg_.tree_max_depth;
        T.valid = true;
        stats_.tree_rebuilt = 1;
    }
    hipLaunchKernelGGL(k_tc_gather, dim3(nblocks(n)), dim3(kNT), 0, stream_, T.d_q, T.d_ord, (int)n, T.d_qp);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tc_leaf_moments, dim3(nblocks(T.nleaves)), dim3(kNT), 0, stream_, T.d_nodes, T.d_leaves,
                       T.nleaves, T.d_pos, T.d_qp, T.d_mom);
    LLGMI_LAUNCH_CHECK();
    for (int d = (int)T.internal_by_depth.size() - 1; d >= 0; --d) {
        const auto od = T.internal_by_depth[(size_t)d];
        if (od.second == 0) continue;
        hipLaunchKernelGGL(k_tc_node_moments, dim3(nblocks(od.second)), dim3(kNT), 0, stream_, T.d_nodes,
                           T.d_internal + od.first, od.second, T.d_mom);
        LLGMI_LAUNCH_CHECK();
    }
    BoxGeom bg{px_, py_, pz_, L.dx, L.dy, L.dz, (double)px_ * 0.5, (double)py_ * 0.5, (double)pz_ * 0.5};
    if (!T.d_targets) {
        // boundary cells of the finest box in 8 x 8 patches of each face,
        // each patch padded to 64 entries (one wave)
        std::vector<int> tg;
        auto face = [&](int nu, int nv, auto &&cell_of) {
            for (int v0 = 0; v0 < nv; v0 += 8)
                for (int u0 = 0; u0 < nu; u0 += 8)
                    for (int l = 0; l < 64; ++l) {
                        const int u = u0 + l % 8, v = v0 + l / 8;
                        tg.push_back(u < nu && v < nv ? cell_of(u, v) : -1);
                    }
        };
        const int nx = px_, ny = py_, nz = pz_;
        auto at = [&](int i, int j, int k) { return (int)(((int64_t)k * ny + j) * nx + i); };
        face(nx, ny, [&](int u, int v) { return at(u, v, 0); });
        face(nx, ny, [&](int u, int v) { return at(u, v, nz - 1); });
        face(nx, nz - 2, [&](int u, int v) { return at(u, 0, v + 1); });
        face(nx, nz - 2, [&](int u, int v) { return at(u, ny - 1, v + 1); });
        face(ny - 2, nz - 2, [&](int u, int v) { return at(0, u + 1, v + 1); });
        face(ny - 2, nz - 2, [&](int u, int v) { return at(nx - 1, u + 1, v + 1); });
        T.ntargets = (int)tg.size();
        dev_alloc(T.d_targets, tg.size());
        LLGMI_HIP(hipMemcpy(T.d_targets, tg.data(), sizeof(int) * tg.size(), hipMemcpyHostToDevice));
    }
    const size_t lds = sizeof(double) * kNT * (size_t)(T.deepest + 1);
    if (lds > 64 * 1024) {
        if (lds > 160 * 1024) throw Error(LLGMI_ERR_UNSUPPORTED, "treecode deeper than the LDS holds");
        LLGMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_tc_eval),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL(k_tc_eval, dim3(nblocks(T.ntargets)), dim3(kNT), lds, stream_, T.d_nodes, T.nnodes, T.d_mom,
                       T.d_pos, T.d_qp, bg, cfg_.tree_theta, T.d_targets, T.ntargets, bc_, L.phi[L.cur]);
    LLGMI_LAUNCH_CHECK();
}

void MGSolver::dipole_bc() {
    Level &L = lev_[0];
    const int k_lo = std::max(1, offz_ - 1), k_hi = std::min(pz_ - 2, offz_ + 1);
    BoxGeom bg{px_, py_, pz_, L.dx, L.dy, L.dz, (double)px_ * 0.5, (double)py_ * 0.5, (double)pz_ * 0.5};
    const double dvol = L.dx * L.dy * L.dz;
    hipLaunchKernelGGL(k_dip_partial, dim3(kRedBlocks), dim3(kNT), 0, stream_, L.rhs, bg, k_lo, k_hi - k_lo + 1,
                       dvol, d_partial_);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dip_final, dim3(1), dim3(64), 0, stream_, d_partial_, kRedBlocks, d_scalars_ + 2);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dip_eval, dim3(nblocks(boundary_count(px_, py_, pz_))), dim3(kNT), 0, stream_,
                       d_scalars_ + 2, bg, bc_, L.phi[L.cur]);
    LLGMI_LAUNCH_CHECK();
}

// update_finest_boundary_bc (:1804-1833): boundary values, then the stamp
void MGSolver::update_bc() {
    ProfScope ps(prof_, LLGMI_K_MG_BC, stream_);
    stats_.n_charges = 0;
    stats_.tree_rebuilt = 0;
    if (cfg_.bc == 0) {
        LLGMI_HIP(hipMemsetAsync(bc_, 0, sizeof(double) * (size_t)px_ * py_ * pz_, stream_));
        stamp(0);
    } else if (cfg_.bc == 1) {
        dipole_bc();
    } else {
        treecode_bc();
    }
}

// second phi buffer of a level (Jacobi / fused red-black sweeps write into it)
void MGSolver::ensure_pair(Level &L) {
    if (L.phi[1]) return;
    const size_t n = (size_t)L.nx * L.ny * L.nz;
    LLGMI_HIP(hipMalloc((void **)&L.phi[1], sizeof(double) * n));
    LLGMI_HIP(hipMemsetAsync(L.phi[1], 0, sizeof(double) * n, stream_));
}

void MGSolver::smooth(int l, int64_t iters) {
    Level &L = lev_[l];
    if (cfg_.smoother == 1) {
        const double sx = 1.0 / (L.dx * L.dx), sy = 1.0 / (L.dy * L.dy), sz = 1.0 / (L.dz * L.dz);
        const double denom = 2.0 * (sx + sy + sz);
        const bool coarsest = l == (int)lev_.size() - 1;
        if (rb_fused_ && l == 0) {
            // finest level: one fused kernel per sweep
            ensure_pair(L);
            const int rbz = std::max(1, std::min(rbz_, L.nz - 2));
            const dim3 grid((unsigned)((L.nx - 2 + RBX - 1) / RBX), (unsigned)((L.ny - 2 + RBY - 1) / RBY),
                            (unsigned)((L.nz - 2 + rbz - 1) / rbz));
            for (int64_t it = 0; it < iters; ++it) {
                hipLaunchKernelGGL(k_mg_rb_sweep, grid, dim3(RBX * RBY), 0, stream_, L.phi[L.cur],
                                   L.phi[1 - L.cur], L.rhs, L.nx, L.ny, L.nz, sx, sy, sz, denom, cfg_.sor_omega,
                                   rbz);
                LLGMI_LAUNCH_CHECK();
                L.cur = 1 - L.cur;
            }
            return;
        }
        if (rb_fused_ && coarsest && l > 0 && L.nz <= TB_MAXZ) {
            // small coarsest level: TBH half-sweeps per launch in LDS
            ensure_pair(L);
            const dim3 grid((unsigned)((L.nx - 2 + TBT - 1) / TBT), (unsigned)((L.ny - 2 + TBT - 1) / TBT));
            const size_t lds = 2 * sizeof(double) * (size_t)L.nz * TBW * TBW;
            int64_t left = 2 * iters;
            int color = 0;
            while (left > 0) {
                const int h = (int)std::min<int64_t>(left, TBH);
                hipLaunchKernelGGL(k_mg_rb_tb, grid, dim3(TB_NT), lds, stream_, L.phi[L.cur], L.phi[1 - L.cur],
                                   L.rhs, L.nx, L.ny, L.nz, sx, sy, sz, denom, cfg_.sor_omega, color, h);
                LLGMI_LAUNCH_CHECK();
                L.cur = 1 - L.cur;
                color = (color + h) & 1;
                left -= h;
            }
            return;
        }
        // other levels: the two half-sweep kernels, in place
        const int64_t n = (int64_t)(L.ny - 2) * (L.nz - 2) * ((L.nx - 1) / 2);
        for (int64_t it = 0; it < iters; ++it)
            for (int color = 0; color < 2; ++color) {
                hipLaunchKernelGGL(k_mg_rb, dim3(nblocks(n)), dim3(kNT), 0, stream_, L.phi[L.cur], L.rhs, L.nx,
                                   L.ny, L.nz, sx, sy, sz, denom, cfg_.sor_omega, color);
                LLGMI_LAUNCH_CHECK();
            }
        return;
    }
    ensure_pair(L);
    const int64_t n = (int64_t)L.nx * L.ny * L.nz;
    for (int64_t it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_mg_wj, dim3(nblocks(n)), dim3(kNT), 0, stream_, L.phi[L.cur], L.phi[1 - L.cur], L.rhs,
                           l == 0 ? bc_ : nullptr, (const DevStencil *)L.d_st, cfg_.omega, L.nx, L.ny, L.nz);
        LLGMI_LAUNCH_CHECK();
        L.cur = 1 - L.cur;
    }
}

// v_cycle (:1877-1945)
void MGSolver::v_cycle(int l) {
    if (l == (int)lev_.size() - 1) {
        smooth(l, 80);
        return;
    }
    smooth(l, cfg_.pre_smooth);
    Level &F = lev_[l], &C = lev_[l + 1];
    const int rx = F.nx / C.nx, ry = F.ny / C.ny, rz = F.nz / C.nz;
    const int64_t nc = (int64_t)(C.nx - 2) * (C.ny - 2) * (C.nz - 2);
    hipLaunchKernelGGL(k_mg_resid_restrict, dim3(nblocks(nc)), dim3(kNT), 0, stream_, F.phi[F.cur], F.rhs,
                       (const DevStencil *)F.d_st, F.nx, F.ny, F.nz, C.rhs, C.phi[C.cur], C.nx, C.ny, C.nz, rx, ry,
                       rz, 1.0 / (double)(rx * ry * rz), F.seven ? 1 : 0);
    LLGMI_LAUNCH_CHECK();
    v_cycle(l + 1);
    const int64_t nf = (int64_t)(F.nx - 2) * (F.ny - 2) * (F.nz - 2);
    hipLaunchKernelGGL(k_mg_prolong, dim3(nblocks(nf)), dim3(kNT), 0, stream_, C.phi[C.cur], C.nx, C.ny, C.nz,
                       F.phi[F.cur], F.nx, F.ny, F.nz, cfg_.prolong == 1 ? 1 : 0);
    LLGMI_LAUNCH_CHECK();
    smooth(l, cfg_.post_smooth);
}

// One V-cycle from a HIP graph: the cycle's ~180 launches (80 coarsest
// sweeps) are captured once per smoothing configuration and replayed, which
// removes the host launch cost between the short coarse-level kernels.  The
// captured work has no memset/memcpy nodes; every buffer it touches exists
// before the capture (the Jacobi target buffers are allocated first).
void MGSolver::v_cycle_graph() {
    if (!use_graphs_) {
        v_cycle(0);
        return;
    }
    if (cfg_.smoother == 0 || rb_fused_)
        for (auto &L : lev_) ensure_pair(L);
    std::string key = std::to_string(cfg_.smoother) + "/" + std::to_string(cfg_.pre_smooth) + "/" +
                      std::to_string(cfg_.post_smooth) + "/";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%a/%a/", cfg_.omega, cfg_.sor_omega);
    key += buf;
    for (auto &L : lev_) key += (char)('0' + L.cur);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        std::vector<int> cur_before;
        for (auto &L : lev_) cur_before.push_back(L.cur);
        hipGraph_t graph = nullptr;
        LLGMI_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            v_cycle(0);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            for (size_t l = 0; l < lev_.size(); ++l) lev_[l].cur = cur_before[l];
            throw;
        }
        LLGMI_HIP(hipStreamEndCapture(stream_, &graph));
        Graph g;
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        LLGMI_HIP(e);
        for (auto &L : lev_) g.cur_after.push_back(L.cur);
        for (size_t l = 0; l < lev_.size(); ++l) lev_[l].cur = cur_before[l];
        it = graphs_.emplace(key, std::move(g)).first;
    }
    LLGMI_HIP(hipGraphLaunch(it->second.exec, stream_));
    for (size_t l = 0; l < lev_.size(); ++l) lev_[l].cur = it->second.cur_after[l];
}

double MGSolver::finest_rhs_max() {
    const int k_lo = std::max(1, offz_ - 1), k_hi = std::min(pz_ - 2, offz_ + 1);
    hipLaunchKernelGGL(k_mg_rhs_absmax, dim3(kRedBlocks), dim3(kNT), 0, stream_, lev_[0].rhs, px_, py_, k_lo,
                       k_hi - k_lo + 1, d_partial_);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mg_final_max, dim3(1), dim3(kNT), 0, stream_, d_partial_, kRedBlocks, d_scalars_);
    LLGMI_LAUNCH_CHECK();
    LLGMI_HIP(hipMemcpyAsync(h_scalars_, d_scalars_, sizeof(double), hipMemcpyDeviceToHost, stream_));
    LLGMI_HIP(hipStreamSynchronize(stream_));
    return h_scalars_[0];
}

// solve_with_timing (:1947-2057)
void MGSolver::phi_copy(double *dev, bool into) {
    Level &L = lev_[0];
    const size_t bytes = sizeof(double) * (size_t)L.nx * L.ny * L.nz;
    if (into) LLGMI_HIP(hipMemcpyAsync(L.phi[L.cur], dev, bytes, hipMemcpyDeviceToDevice, stream_));
    else LLGMI_HIP(hipMemcpyAsync(dev, L.phi[L.cur], bytes, hipMemcpyDeviceToDevice, stream_));
    LLGMI_HIP(hipStreamSynchronize(stream_));
}

void MGSolver::solve() {
    Level &L = lev_[0];
    const size_t n0 = (size_t)L.nx * L.ny * L.nz;
    if (!cfg_.warm_start) LLGMI_HIP(hipMemsetAsync(L.phi[L.cur], 0, sizeof(double) * n0, stream_));
    update_bc();   // ends with the boundary stamped (enforce_dirichlet)
    ProfScope ps(prof_, LLGMI_K_MG, stream_);
    const bool use_tol = cfg_.has_tol_abs || cfg_.has_tol_rel;
    const double rhs_max = (use_tol && cfg_.has_tol_rel) ? finest_rhs_max() : 0.0;
    double tol_target = 0.0;
    if (use_tol) {
        if (cfg_.has_tol_abs) tol_target = std::fmax(tol_target, std::fmax(cfg_.tol_abs, 0.0));
        if (cfg_.has_tol_rel) tol_target = std::fmax(tol_target, std::fmax(std::fmax(cfg_.tol_rel, 0.0) * rhs_max, 0.0));
    }
    const int64_t min_cycles = use_tol ? std::max<int64_t>(cfg_.v_cycles, 1) : 0;
    const int64_t max_cycles = use_tol ? std::max<int64_t>(std::max<int64_t>(cfg_.v_cycles_max, min_cycles), 1)
                                       : std::max<int64_t>(cfg_.v_cycles, 1);
    stats_.rhs_max = rhs_max;
    stats_.tol_target = tol_target;
    stats_.last_residual = -1.0;
    stats_.diverged = 0;
    stats_.cycles = 0;
    const int64_t n_nan = std::min<int64_t>(2 * (int64_t)L.nx * L.ny, (int64_t)n0);
    for (int64_t iter = 0; iter < max_cycles; ++iter) {
        v_cycle_graph();
        // enforce_dirichlet: the smoothers leave the boundary cells at bc
        stats_.cycles = iter + 1;
        const bool nan_check = iter % 2 == 1 || iter + 1 == max_cycles;
        const bool res_check = use_tol && iter + 1 >= min_cycles;
        if (!nan_check && !res_check) continue;
        if (nan_check) {
            LLGMI_HIP(hipMemsetAsync(d_scalars_ + 1, 0, sizeof(double), stream_));
            hipLaunchKernelGGL(k_mg_nan, dim3(nblocks(n_nan)), dim3(kNT), 0, stream_, L.phi[L.cur], n_nan,
                               d_scalars_ + 1);
            LLGMI_LAUNCH_CHECK();
        }
        if (res_check) {
            hipLaunchKernelGGL(k_mg_resid_max, dim3(kRedBlocks), dim3(kNT), 0, stream_, L.phi[L.cur], L.rhs,
                               (const DevStencil *)L.d_st, L.nx, L.ny, L.nz, L.seven ? 1 : 0, d_partial_);
            LLGMI_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_mg_final_max, dim3(1), dim3(kNT), 0, stream_, d_partial_, kRedBlocks, d_scalars_);
            LLGMI_LAUNCH_CHECK();
        }
        LLGMI_HIP(hipMemcpyAsync(h_scalars_, d_scalars_, sizeof(double) * 2, hipMemcpyDeviceToHost, stream_));
        LLGMI_HIP(hipStreamSynchronize(stream_));
        if (nan_check && h_scalars_[1] != 0.0) {
            LLGMI_HIP(hipMemsetAsync(L.phi[L.cur], 0, sizeof(double) * n0, stream_));
            stamp(0);
            stats_.diverged = 1;
            break;
        }
        if (res_check) {
            stats_.last_residual = h_scalars_[0];
            if (h_scalars_[0] <= tol_target) break;
        }
    }
}

template <typename T>
void MGSolver::gradient(const T *m, T *bd) {
    const Level &L = lev_[0];
    const int64_t n = g_.nx * g_.ny;
    hipLaunchKernelGGL(k_mg_gradient<T>, dim3(nblocks(n)), dim3(kNT), 0, stream_, L.phi[L.cur], px_, py_, pz_, L.dx,
                       L.dy, L.dz, offx_, offy_, offz_, (int)g_.nx, (int)g_.ny, m, bd);
    LLGMI_LAUNCH_CHECK();
}

// DemagPoissonMG::add_field (:2094-2098): rhs, solve, gradient; with the
// hybrid's Gaussian screening of the rhs (:2518-2520) when sigma > 0.
template <typename T>
void MGSolver::demag_field(const T *m, double ms, T *bd, double sigma, bool masked) {
    {
        ProfScope ps(prof_, LLGMI_K_MG, stream_);
        build_rhs(m, std::is_same<T, float>::value, ms);
        if (sigma > 0.0) screen_rhs(sigma);
    }
    solve();
    ProfScope ps(prof_, LLGMI_K_MG, stream_);
    gradient(masked ? m : (const T *)nullptr, bd);
}

template void MGSolver::demag_field<float>(const float *, double, float *, double, bool);
template void MGSolver::demag_field<double>(const double *, double, double *, double, bool);

// gaussian_kernel_1d (mg_kernels.rs:580-599) on the host, the two passes on the
// finest level with the second phi buffer as scratch (free between solves).
void MGSolver::screen_rhs(double sigma) {
    if (!(sigma > 0.0)) return;
    const int64_t r = (int64_t)std::ceil(3.0 * sigma);
    if (r <= 0) return;
    Level &L = lev_[0];
    if (L.nx < 3 || L.ny < 3 || L.nz < 3) return;
    std::vector<double> &w = screen_w_;   // kept alive for the asynchronous upload
    w.assign((size_t)(2 * r + 1), 0.0);
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    for (int64_t q = -r; q <= r; ++q) {
        const double x = (double)q;
        w[(size_t)(q + r)] = std::exp(-x * x * inv2s2);
    }
    double sum = 0.0;
    for (double v : w) sum += v;
    if (sum > 0.0)
        for (double &v : w) v /= sum;
    if (!d_screen_w_ || screen_cap_ < w.size()) {
        if (d_screen_w_) (void)hipFree(d_screen_w_);
        d_screen_w_ = nullptr;
        LLGMI_HIP(hipMalloc(&d_screen_w_, sizeof(double) * w.size()));
        screen_cap_ = w.size();
    }
    LLGMI_HIP(hipMemcpyAsync(d_screen_w_, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, stream_));
    ensure_pair(L);
    double *tmp = L.phi[1 - L.cur];
    const int64_t n = (int64_t)L.nx * L.ny * L.nz;
    hipLaunchKernelGGL(k_mg_screen<0>, dim3(nblocks(n)), dim3(kNT), 0, stream_, L.rhs, tmp, d_screen_w_, (int)r,
                       L.nx, L.ny, L.nz);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mg_screen<1>, dim3(nblocks(n)), dim3(kNT), 0, stream_, tmp, L.rhs, d_screen_w_, (int)r,
                       L.nx, L.ny, L.nz);
    LLGMI_LAUNCH_CHECK();
    // the scratch buffer is the Jacobi / fused-sweep target: its contents are
    // overwritten before they are read there
}

template <typename T>
void launch_hybrid_correction(const double *dk, int r, const T *m, T *b, double ms, double bscale, int bmargin,
                              double *scale_buf, int nx, int ny, hipStream_t s) {
    const int64_t n = (int64_t)nx * ny;
    const double *scale = nullptr;
    if (bscale > 0.0) {
        hipLaunchKernelGGL(k_hyb_scale<T>, dim3(nblocks(n)), dim3(kNT), 0, s, m, nx, ny, r, bmargin, bscale,
                           scale_buf);
        LLGMI_LAUNCH_CHECK();
        scale = scale_buf;
    }
    hipLaunchKernelGGL(k_hyb_correct<T>, dim3(nblocks(n)), dim3(kNT), 0, s, dk, r, m, b, ms, scale, nx, ny);
    LLGMI_LAUNCH_CHECK();
}
template void launch_hybrid_correction<float>(const double *, int, const float *, float *, double, double, int,
                                              double *, int, int, hipStream_t);
template void launch_hybrid_correction<double>(const double *, int, const double *, double *, double, double, int,
                                               double *, int, int, hipStream_t);

}  // namespace llgmi

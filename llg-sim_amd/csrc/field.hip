// Local effective-field kernel (Zeeman, exchange, uniaxial anisotropy,
// interfacial DMI, optional demag addend, optional LLG torque) and the RK45
// stage / error-norm kernels.
//
// Field terms follow llg-sim's effective_field/ (zeeman.rs:6-11,
// exchange.rs:45-222, anisotropy.rs:10-25, dmi.rs:67-209, summed in the order
// of build_h_eff_masked, mod.rs:48-75); the torque is llg_rhs (llg.rs:20-32);
// the stage combination and error norm are those of
// step_llg_rk45_recompute_field_masked_adaptive (llg.rs:869-997).
//
// One thread per cell, SoA component-major arrays; neighbour reads hit the
// same or adjacent cache lines (x) or a row already pulled into L2 by the
// neighbouring block row (y), so the kernel streams m once and B once.
#include "common.hpp"
#include "internal.hpp"
#include "local_cell.hpp"

namespace llgmi {

template <typename T, bool GEOM, bool DEMAG, bool TORQUE>
__global__ void __launch_bounds__(256)
k_local_field(LocalArgs<T> a, const T *__restrict__ m, const T *__restrict__ bd,
              const uint8_t *__restrict__ geom, T *__restrict__ b, T *__restrict__ kout) {
    // neighbouring rows share one XCD's L2 (their halo rows are reused there)
    const int lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int j = lb / gridDim.x;
    const int i = (lb - j * gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= a.nx) return;
    T bd3[3] = {T(0), T(0), T(0)};
    if (DEMAG) {
        const int64_t n = (int64_t)a.nx * a.ny, c = (int64_t)j * a.nx + i;
        bd3[0] = bd[c];
        bd3[1] = bd[n + c];
        bd3[2] = bd[2 * n + c];
    }
    local_cell<T, GEOM, DEMAG, TORQUE>(a, m, geom, i, j, bd3, b, kout);
}

template <typename T>
void launch_local_field(const LocalArgs<T> &a, const T *m, const T *bd, const uint8_t *geom,
                        T *b, T *k, hipStream_t s) {
    dim3 blk(256), grid((unsigned)((a.nx + 255) / 256), (unsigned)a.ny);
    const bool g = geom != nullptr, d = bd != nullptr, t = a.torque != 0;
#define LLGMI_LF(G, D, Q) \
    hipLaunchKernelGGL((k_local_field<T, G, D, Q>), grid, blk, 0, s, a, m, bd, geom, b, k)
    if (g) {
        if (d) { if (t) LLGMI_LF(true, true, true); else LLGMI_LF(true, true, false); }
        else { if (t) LLGMI_LF(true, false, true); else LLGMI_LF(true, false, false); }
    } else {
        if (d) { if (t) LLGMI_LF(false, true, true); else LLGMI_LF(false, true, false); }
        else { if (t) LLGMI_LF(false, false, true); else LLGMI_LF(false, false, false); }
    }
#undef LLGMI_LF
    LLGMI_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's AoS f64 and device SoA
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_aos_to_soa(const double *__restrict__ aos, T *__restrict__ soa, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    soa[i] = T(aos[3 * i]);
    soa[n + i] = T(aos[3 * i + 1]);
    soa[2 * n + i] = T(aos[3 * i + 2]);
}
template <typename T>
__global__ void k_soa_to_aos(const T *__restrict__ soa, double *__restrict__ aos, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] = (double)soa[i];
    aos[3 * i + 1] = (double)soa[n + i];
    aos[3 * i + 2] = (double)soa[2 * n + i];
}
template <typename T>
__global__ void k_soa_add_to_aos(const T *__restrict__ soa, double *__restrict__ aos, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] += (double)soa[i];
    aos[3 * i + 1] += (double)soa[n + i];
    aos[3 * i + 2] += (double)soa[2 * n + i];
}

template <typename T>
void launch_aos_to_soa(const double *aos, T *soa, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_aos_to_soa<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, aos, soa, n);
    LLGMI_LAUNCH_CHECK();
}
template <typename T>
void launch_soa_to_aos(const T *soa, double *aos, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_soa_to_aos<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, soa, aos, n);
    LLGMI_LAUNCH_CHECK();
}
template <typename T>
void launch_soa_add_to_aos(const T *soa, double *aos, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_soa_add_to_aos<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, soa, aos, n);
    LLGMI_LAUNCH_CHECK();
}

template <typename T>
__global__ void k_accumulate(const T *__restrict__ src, T *__restrict__ dst, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] += src[i];
}
template <typename T>
void launch_soa_accumulate(const T *src, T *dst, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL((k_accumulate<T>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
    LLGMI_LAUNCH_CHECK();
}

// m = 0 on cells whose keep byte is 0 (SoA, n cells)
template <typename T>
__global__ void k_mask_cells(T *__restrict__ m, const uint8_t *__restrict__ keep, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !keep[i]) {
        m[i] = T(0);
        m[n + i] = T(0);
        m[2 * n + i] = T(0);
    }
}
template <typename T>
void launch_mask_cells(T *m, const uint8_t *keep, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_mask_cells<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, keep, n);
    LLGMI_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// RK45 stage state: v = m; v = v + c_q k_q (q in order); normalize (zero -> +z)
// ---------------------------------------------------------------------------
// cell i: v = m + c_q k_q (q in order), normalised, into out
template <typename T>
__device__ __forceinline__ void rk_stage_cell(const T *__restrict__ m, const T *const *k, const T *c, int nk,
                                              T *__restrict__ out, int64_t i, int64_t n) {
    T v[3] = {m[i], m[n + i], m[2 * n + i]};
    stage_combine<T>(v, k, c, nk, nullptr, i, n);
    out[i] = v[0];
    out[n + i] = v[1];
    out[2 * n + i] = v[2];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_rk_stage(const T *__restrict__ m, StageArgs<T> a, T *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T c[6];
    for (int q = 0; q < a.nk; ++q) c[q] = a.dt ? T(*a.dt * a.a[q]) : a.c[q];
    rk_stage_cell<T>(m, a.k, c, a.nk, out, i, n);
}

template <typename T>
void launch_rk_stage(const T *m, const StageArgs<T> &a, T *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_rk_stage<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, a, out, n);
    LLGMI_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Error norm (llg.rs:948-997), FP64.  reference form: |tau5 - tau4| with the
// two weighted sums formed separately (FP64 contexts, bit-for-bit the
// reference's arithmetic order); combined form: |sum_j (b_j - bh_j) k_j| with
// FP64 coefficients (FP32 contexts, where forming tau5 and tau4 from FP32
// torques would cancel the estimate away).
// ---------------------------------------------------------------------------
struct DP {
    static constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                            B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
    static constexpr double BH1 = 5179.0 / 57600.0, BH3 = 7571.0 / 16695.0, BH4 = 393.0 / 640.0,
                            BH5 = -92097.0 / 339200.0, BH6 = 187.0 / 2100.0, BH7 = 1.0 / 40.0;
};

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// max over the three components of cell t of |error estimate|
template <typename T, bool COMBINED>
__device__ __forceinline__ double rk_err_cell(const T *__restrict__ k1, const T *__restrict__ k3,
                                              const T *__restrict__ k4, const T *__restrict__ k5,
                                              const T *__restrict__ k6, const T *__restrict__ k7, int64_t t,
                                              int64_t n) {
#pragma clang fp contract(off)   // the same rounding in every kernel that inlines it
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        int64_t o = q * n + t;
        double a1 = k1[o], a3 = k3[o], a4 = k4[o], a5 = k5[o], a6 = k6[o], a7 = k7[o];
        double e;
        if (COMBINED) {
            e = (DP::B1 - DP::BH1) * a1 + (DP::B3 - DP::BH3) * a3 + (DP::B4 - DP::BH4) * a4 +
                (DP::B5 - DP::BH5) * a5 + (DP::B6 - DP::BH6) * a6 - DP::BH7 * a7;
        } else {
            double th = DP::B1 * a1 + DP::B3 * a3 + DP::B4 * a4 + DP::B5 * a5 + DP::B6 * a6;
            double tl = DP::BH1 * a1 + DP::BH3 * a3 + DP::BH4 * a4 + DP::BH5 * a5 + DP::BH6 * a6 + DP::BH7 * a7;
            e = th - tl;
        }
        d = q == 0 ? fabs(e) : fmax(d, fabs(e));
    }
    return d;
}

template <typename T, bool COMBINED>
__global__ void __launch_bounds__(256)
k_rk_err(const T *__restrict__ k1, const T *__restrict__ k3, const T *__restrict__ k4,
         const T *__restrict__ k5, const T *__restrict__ k6, const T *__restrict__ k7, int64_t n,
         double *__restrict__ partial) {
    __shared__ double red[4];
    double local = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double d = rk_err_cell<T, COMBINED>(k1, k3, k4, k5, k6, k7, t, n);
        if (d > local) local = d;
    }
    local = wave_max(local);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[w]);
        partial[blockIdx.x] = v;
    }
}

// One film's controller (llg.rs:999-1025, the same decision as kb_control)
// from the attempt's max |err|: attempt a's outputs, the next step size into
// *c.dt and the decision into *c.flag.
__device__ __forceinline__ void rk_control_one(double err, const RkControl &c) {
    const double dt0 = *c.dt;
    const double eps = err * dt0;
    double dt_next = eps > 0.0 ? dt0 * c.headroom * pow(c.max_err / eps, 0.2) : dt0 * 2.0;
    if (dt_next < c.dt_min) dt_next = c.dt_min;
    if (dt_next > c.dt_max) dt_next = c.dt_max;
    const int accept = (eps <= c.max_err) || (dt0 <= c.dt_min * 1.0000000001);
    c.o_eps[c.a] = eps;
    c.o_acc[c.a] = accept;
    c.o_dtu[c.a] = dt0;
    *c.dt = dt_next;
    *c.flag = accept;
}

// max of the partial rows into out[0]; with c.dt set, then the controller
__global__ void __launch_bounds__(1024) k_max_reduce(const double *__restrict__ partial, int np,
                                                     double *__restrict__ out, RkControl c) {
    __shared__ double red[16];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v = fmax(v, partial[i]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
        out[0] = r;
        if (c.dt) rk_control_one(r, c);
    }
}

template <typename T>
void launch_rk_err(const T *k1, const T *k3, const T *k4, const T *k5, const T *k6, const T *k7,
                   int64_t n, bool combined, double *partial, double *out, hipStream_t s, const RkControl *ctl) {
    int64_t want = (n + 255) / 256;
    int nb = (int)(want < kErrBlocks ? (want > 0 ? want : 1) : kErrBlocks);
    if (combined)
        hipLaunchKernelGGL((k_rk_err<T, true>), dim3(nb), dim3(256), 0, s, k1, k3, k4, k5, k6, k7, n, partial);
    else
        hipLaunchKernelGGL((k_rk_err<T, false>), dim3(nb), dim3(256), 0, s, k1, k3, k4, k5, k6, k7, n, partial);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(1024), 0, s, partial, nb, out, ctl ? *ctl : RkControl{});
    LLGMI_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Batched element-wise kernels: entry blockIdx.x = (film, block of kBatchCells
// cells of that film).
// ---------------------------------------------------------------------------
template <typename T, bool DEMAG>
__global__ void __launch_bounds__(kBatchCells)
kb_local(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int which, int ks, int stage1) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (stage1 && f.skip1) return;
    const int64_t n = (int64_t)f.nx * f.ny;
    const int64_t c = (int64_t)e.blk * kBatchCells + threadIdx.x;
    if (c >= n) return;
    const int j = (int)(c / f.nx), i = (int)(c - (int64_t)j * f.nx);
    T bd3[3] = {T(0), T(0), T(0)};
    if (DEMAG) {
        bd3[0] = f.bd[c];
        bd3[1] = f.bd[n + c];
        bd3[2] = f.bd[2 * n + c];
    }
    local_cell<T, false, DEMAG, true>(f.la, f.in[which], nullptr, i, j, bd3, f.b, f.k[ks]);
}

// Dormand-Prince stage coefficients a_sj (llg.rs:813-846) and the 5th-order
// weights of m7, with the k indices they multiply
struct DPStage {
    int nk, kidx[5];
    double a[5];
};
__constant__ DPStage kDP[8] = {
    {0, {0}, {0}}, {0, {0}, {0}},
    {1, {1}, {1.0 / 5.0}},
    {2, {1, 2}, {3.0 / 40.0, 9.0 / 40.0}},
    {3, {1, 2, 3}, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0}},
    {4, {1, 2, 3, 4}, {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0}},
    {5, {1, 2, 3, 4, 5}, {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}},
    {5, {1, 3, 4, 5, 6}, {35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}},
};

template <typename T>
__global__ void __launch_bounds__(kBatchCells)
kb_stage(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int stage) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    const int64_t n = (int64_t)f.nx * f.ny;
    const int64_t c = (int64_t)e.blk * kBatchCells + threadIdx.x;
    if (c >= n) return;
    const DPStage &st = kDP[stage];
    const T *k[6];
    T co[6];
    for (int q = 0; q < st.nk; ++q) {
        k[q] = f.k[st.kidx[q]];
        co[q] = T(f.dt * st.a[q]);
    }
    rk_stage_cell<T>(f.m, k, co, st.nk, stage == 7 ? f.m7 : f.stage, c, n);
}

template <typename T, bool COMBINED>
__global__ void __launch_bounds__(kBatchCells)
kb_err(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, double *__restrict__ partial) {
    __shared__ double red[kBatchCells / 64];
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    const int64_t n = (int64_t)f.nx * f.ny;
    const int64_t c = (int64_t)e.blk * kBatchCells + threadIdx.x;
    double local = c < n ? rk_err_cell<T, COMBINED>(f.k[1], f.k[3], f.k[4], f.k[5], f.k[6], f.k[7], c, n) : 0.0;
    local = wave_max(local);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < kBatchCells / 64; ++w) v = fmax(v, red[w]);
        partial[blockIdx.x] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
kb_err_reduce(const BFilm<T> *__restrict__ F, const double *__restrict__ partial, double *__restrict__ err) {
    __shared__ double red[4];
    const BFilm<T> &f = F[blockIdx.x];
    double v = 0.0;
    for (int q = threadIdx.x; q < f.ne; q += blockDim.x) v = fmax(v, partial[f.e0 + q]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) err[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

template <typename T>
__global__ void __launch_bounds__(kBatchCells)
kb_accept(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (!f.accept) return;
    const int64_t n = (int64_t)f.nx * f.ny;
    const int64_t c = (int64_t)e.blk * kBatchCells + threadIdx.x;
    if (c >= n) return;
    f.m[c] = f.m7[c];
    f.m[n + c] = f.m7[n + c];
    f.m[2 * n + c] = f.m7[2 * n + c];
}

// After an accepted attempt: m = m7 and k1 = k7 (the next attempt's first
// slope, FSAL); nothing after a rejected one (m and k1 = f(m) stay).
template <typename T>
__global__ void __launch_bounds__(256)
k_rk_accept(const int *__restrict__ flag, T *__restrict__ m, const T *__restrict__ m7, T *__restrict__ k1,
            const T *__restrict__ k7, int64_t n3) {
    if (!*flag) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        m[i] = m7[i];
        k1[i] = k7[i];
    }
}

// Stage 2 of an attempt inside a device-controlled run: after an accepted
// attempt (*flag set) the base state and first slope are m7 and k7, which are
// also copied to m and k1 (what k_rk_accept does), else m and k1; then the
// stage state normalize(base + c k) goes to out.  a: stage 2's arguments
// (one slope).
template <typename T>
__global__ void __launch_bounds__(256)
k_rk_stage2_accept(const int *__restrict__ flag, T *__restrict__ m, const T *__restrict__ m7, T *__restrict__ k1,
                   const T *__restrict__ k7, StageArgs<T> a, T *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool acc = *flag != 0;
    const T *mb = acc ? m7 : m;
    const T *kb = acc ? k7 : k1;
    const T base[3] = {mb[i], mb[n + i], mb[2 * n + i]};
    const T kk[3] = {kb[i], kb[n + i], kb[2 * n + i]};
    const T c0 = a.dt ? T(*a.dt * a.a[0]) : a.c[0];
    T v[3] = {base[0], base[1], base[2]};
    stage_combine<T>(v, nullptr, &c0, 1, kk, i, n);
    out[i] = v[0];
    out[n + i] = v[1];
    out[2 * n + i] = v[2];
    if (acc) {
        m[i] = base[0];
        m[n + i] = base[1];
        m[2 * n + i] = base[2];
        k1[i] = kk[0];
        k1[n + i] = kk[1];
        k1[2 * n + i] = kk[2];
    }
}

template <typename T>
void launch_rk_stage2_accept(const int *flag, T *m, const T *m7, T *k1, const T *k7, const StageArgs<T> &a,
                             T *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_rk_stage2_accept<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, flag, m, m7, k1,
                       k7, a, out, n);
    LLGMI_LAUNCH_CHECK();
}

template <typename T>
void launch_rk_accept(const int *flag, T *m, const T *m7, T *k1, const T *k7, int64_t n, hipStream_t s) {
    const int64_t n3 = 3 * n;
    const int64_t want = (n3 + 255) / 256;
    const unsigned nb = (unsigned)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL((k_rk_accept<T>), dim3(nb), dim3(256), 0, s, flag, m, m7, k1, k7, n3);
    LLGMI_LAUNCH_CHECK();
}

// The controller of step_llg_rk45_recompute_field_masked_adaptive
// (llg.rs:999-1025) on the device, one thread per film, FP64: from the
// attempt's max |err| it forms eps and the next dt, records attempt a's eps,
// decision and step size, and prepares the film's next attempt (dt, FSAL:
// on accept k7 becomes k1 and stage 1 is skipped; the accepted state is
// copied by kb_accept, which runs next).
template <typename T>
__global__ void __launch_bounds__(64)
kb_control(BFilm<T> *__restrict__ F, int nfilms, const double *__restrict__ err, double max_err, double headroom,
           double dt_min, double dt_max, int a, double *__restrict__ o_eps, int *__restrict__ o_acc,
           double *__restrict__ o_dtu) {
    const int f = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (f >= nfilms) return;
    BFilm<T> &r = F[f];
    const double dt0 = r.dt;
    const double eps = err[f] * dt0;
    double dt_next = eps > 0.0 ? dt0 * headroom * pow(max_err / eps, 0.2) : dt0 * 2.0;
    if (dt_next < dt_min) dt_next = dt_min;
    if (dt_next > dt_max) dt_next = dt_max;
    const int accept = (eps <= max_err) || (dt0 <= dt_min * 1.0000000001);
    const int64_t o = (int64_t)a * nfilms + f;
    o_eps[o] = eps;
    o_acc[o] = accept;
    o_dtu[o] = dt0;
    r.dt = dt_next;
    r.accept = accept;
    r.skip1 = accept;
    if (accept) {
        T *t = r.k[1];
        r.k[1] = r.k[7];
        r.k[7] = t;
    }
}

template <typename T>
void launch_batch_control(BFilm<T> *F, int nfilms, const double *err, double max_err, double headroom, double dt_min,
                          double dt_max, int a, double *o_eps, int *o_acc, double *o_dtu, hipStream_t s) {
    if (nfilms <= 0) return;
    hipLaunchKernelGGL((kb_control<T>), dim3((unsigned)((nfilms + 63) / 64)), dim3(64), 0, s, F, nfilms, err, max_err,
                       headroom, dt_min, dt_max, a, o_eps, o_acc, o_dtu);
    LLGMI_LAUNCH_CHECK();
}

template <typename T>
void launch_batch_local(const BFilm<T> *F, const BEntry *tab, int nent, int demag, int which, int ks, int stage1,
                        hipStream_t s) {
    if (nent <= 0) return;
    if (demag)
        hipLaunchKernelGGL((kb_local<T, true>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab, which, ks, stage1);
    else
        hipLaunchKernelGGL((kb_local<T, false>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab, which, ks, stage1);
    LLGMI_LAUNCH_CHECK();
}
template <typename T>
void launch_batch_stage(const BFilm<T> *F, const BEntry *tab, int nent, int stage, hipStream_t s) {
    if (nent <= 0) return;
    hipLaunchKernelGGL((kb_stage<T>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab, stage);
    LLGMI_LAUNCH_CHECK();
}
template <typename T>
void launch_batch_err(const BFilm<T> *F, const BEntry *tab, int nent, int nfilms, double *partial, double *err,
                      hipStream_t s) {
    if (nent <= 0) return;
    if (sizeof(T) == sizeof(float))
        hipLaunchKernelGGL((kb_err<T, true>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab, partial);
    else
        hipLaunchKernelGGL((kb_err<T, false>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab, partial);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL((kb_err_reduce<T>), dim3((unsigned)nfilms), dim3(256), 0, s, F, partial, err);
    LLGMI_LAUNCH_CHECK();
}
template <typename T>
void launch_batch_accept(const BFilm<T> *F, const BEntry *tab, int nent, hipStream_t s) {
    if (nent <= 0) return;
    hipLaunchKernelGGL((kb_accept<T>), dim3((unsigned)nent), dim3(kBatchCells), 0, s, F, tab);
    LLGMI_LAUNCH_CHECK();
}

#define LLGMI_INST(T)                                                                        \
    template void launch_batch_local<T>(const BFilm<T> *, const BEntry *, int, int, int, int, int, \
                                        hipStream_t);                                        \
    template void launch_batch_stage<T>(const BFilm<T> *, const BEntry *, int, int, hipStream_t); \
    template void launch_batch_err<T>(const BFilm<T> *, const BEntry *, int, int, double *,  \
                                      double *, hipStream_t);                                \
    template void launch_batch_accept<T>(const BFilm<T> *, const BEntry *, int, hipStream_t); \
    template void launch_rk_accept<T>(const int *, T *, const T *, T *, const T *, int64_t, hipStream_t); \
    template void launch_rk_stage2_accept<T>(const int *, T *, const T *, T *, const T *, const StageArgs<T> &, \
                                             T *, int64_t, hipStream_t);                                        \
    template void launch_batch_control<T>(BFilm<T> *, int, const double *, double, double, double, double, int, \
                                          double *, int *, double *, hipStream_t);             \
    template void launch_local_field<T>(const LocalArgs<T> &, const T *, const T *,          \
                                        const uint8_t *, T *, T *, hipStream_t);             \
    template void launch_aos_to_soa<T>(const double *, T *, int64_t, hipStream_t);           \
    template void launch_soa_to_aos<T>(const T *, double *, int64_t, hipStream_t);           \
    template void launch_soa_add_to_aos<T>(const T *, double *, int64_t, hipStream_t);       \
    template void launch_soa_accumulate<T>(const T *, T *, int64_t, hipStream_t);            \
    template void launch_mask_cells<T>(T *, const uint8_t *, int64_t, hipStream_t);          \
    template void launch_rk_stage<T>(const T *, const StageArgs<T> &, T *, int64_t, hipStream_t); \
    template void launch_rk_err<T>(const T *, const T *, const T *, const T *, const T *,    \
                                   const T *, int64_t, bool, double *, double *, hipStream_t,     \
                                   const RkControl *);
LLGMI_INST(float)
LLGMI_INST(double)
#undef LLGMI_INST

}  // namespace llgmi

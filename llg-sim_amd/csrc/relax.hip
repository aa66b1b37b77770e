// Kernels of the relaxation path: the error norm of the relax-mode
// Bogacki-Shampine RK23 step, the energy breakdown and the torque metrics.
//
// RK23 error (llg.rs:399-420): eps = max over cells and components of
// |m4 - m2nd|, with m4 the normalised third-order state and m2nd the
// normalised embedded second-order state.  Energy (energy.rs:80-283): the
// reference's per-cell terms, summed per block and then over blocks in a fixed
// order (deterministic for a given grid).  Torque metrics
// (relax.rs:144-190): max, sum and sum of squares of |m x B|.
//
// All three are single-pass streaming reductions over m (and k or B): HBM
// bound, one read of each operand.  Arithmetic is FP64 for either storage
// precision.
#include "common.hpp"
#include "internal.hpp"

namespace llgmi {

namespace {

__device__ __forceinline__ double wmax(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wsum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R>
__device__ __forceinline__ void norm3(R &v0, R &v1, R &v2) {
    R n2 = v0 * v0 + v1 * v1 + v2 * v2;
    if (n2 == R(0)) {
        v0 = R(0); v1 = R(0); v2 = R(1);
    } else {
        R inv = R(1) / sqrt(n2);
        v0 *= inv; v1 *= inv; v2 *= inv;
    }
}

// RECOMPUTE: form the third-order state in FP64 from m, k1..k3 (FP32 storage,
// whose rounding of m4 would otherwise floor the estimate near 1e-7);
// otherwise read the stored m4 (FP64 storage: the same values).
template <typename T, bool RECOMPUTE>
__global__ void __launch_bounds__(256)
k_rk23_err(const T *__restrict__ m, const T *__restrict__ k1, const T *__restrict__ k2,
           const T *__restrict__ k3, const T *__restrict__ k4, const T *__restrict__ m4,
           Rk23ErrCoef c, int64_t n, double *__restrict__ partial) {
    __shared__ double red[4];
    double local = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        double v[3], w[3];   // third-order and embedded second-order states
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int64_t o = q * n + t;
            const double x1 = k1[o], x2 = k2[o], x3 = k3[o], x4 = k4[o];
            double y = (double)m[o];
            y = y + c.bh[0] * x1;
            y = y + c.bh[1] * x2;
            y = y + c.bh[2] * x3;
            y = y + c.bh[3] * x4;
            w[q] = y;
            if (RECOMPUTE) {
                double z = (double)m[o];
                z = z + c.b[0] * x1;
                z = z + c.b[1] * x2;
                z = z + c.b[2] * x3;
                v[q] = z;
            } else {
                v[q] = (double)m4[o];
            }
        }
        norm3(w[0], w[1], w[2]);
        if (RECOMPUTE) norm3(v[0], v[1], v[2]);
        const double d = fmax(fmax(fabs(v[0] - w[0]), fabs(v[1] - w[1])), fabs(v[2] - w[2]));
        if (d > local) local = d;
    }
    local = wmax(local);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, red[q]);
        partial[blockIdx.x] = r;
    }
}

// Reduces np rows of `width` values: out[w] = op over rows (0 max, 1 sum),
// rows visited in a fixed order.
__global__ void __launch_bounds__(1024)
k_rows_reduce(const double *__restrict__ partial, int np, int width, int op_mask,
              double *__restrict__ out) {
    __shared__ double red[16];
    for (int w = 0; w < width; ++w) {
        const bool is_max = (op_mask >> w) & 1;
        double v = 0.0;
        for (int r = threadIdx.x; r < np; r += blockDim.x) {
            double x = partial[(int64_t)r * width + w];
            v = is_max ? fmax(v, x) : v + x;
        }
        v = is_max ? wmax(v) : wsum(v);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = red[0];
            for (int q = 1; q < (int)(blockDim.x >> 6); ++q) s = is_max ? fmax(s, red[q]) : s + red[q];
            out[w] = s;
        }
        __syncthreads();
    }
}

// ghost of energy.rs:36-64 (the y / x component gets + n*d*0)
__device__ __forceinline__ void eg_x(const double *m, double s, double eta, double dx, double *o) {
    o[0] = m[0] + s * dx * (-eta * m[2]);
    o[1] = m[1] + s * dx * 0.0;
    o[2] = m[2] + s * dx * (eta * m[0]);
}
__device__ __forceinline__ void eg_y(const double *m, double s, double eta, double dy, double *o) {
    o[0] = m[0] + s * dy * 0.0;
    o[1] = m[1] + s * dy * (-eta * m[2]);
    o[2] = m[2] + s * dy * (eta * m[1]);
}

template <typename T, bool GEOM, bool DEMAG>
__global__ void __launch_bounds__(256)
k_energy(EnergyArgs a, const T *__restrict__ m, const T *__restrict__ bd,
         const uint8_t *__restrict__ geom, double *__restrict__ partial) {
    __shared__ double red[5][4];
    const int64_t n = (int64_t)a.nx * a.ny;
    const T *mx = m, *my = m + n, *mz = m + 2 * n;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        if (GEOM && !geom[c]) continue;
        const int j = (int)(c / a.nx), i = (int)(c - (int64_t)j * a.nx);
        const double m0[3] = {(double)mx[c], (double)my[c], (double)mz[c]};
        if (a.aex != 0.0) {
            if (i + 1 < a.nx && (!GEOM || geom[c + 1])) {
                double d0 = (double)mx[c + 1] - m0[0], d1 = (double)my[c + 1] - m0[1],
                       d2 = (double)mz[c + 1] - m0[2];
                double sq = d0 * d0 + d1 * d1 + d2 * d2;
                acc[0] += a.aex * (sq / (a.dx * a.dx)) * a.v;
            }
            if (j + 1 < a.ny && (!GEOM || geom[c + a.nx])) {
                const int64_t u = c + a.nx;
                double d0 = (double)mx[u] - m0[0], d1 = (double)my[u] - m0[1], d2 = (double)mz[u] - m0[2];
                double sq = d0 * d0 + d1 * d1 + d2 * d2;
                acc[0] += a.aex * (sq / (a.dy * a.dy)) * a.v;
            }
        }
        if (a.ku != 0.0) {
            double mu = m0[0] * a.u[0] + m0[1] * a.u[1] + m0[2] * a.u[2];
            acc[1] += a.ku * (1.0 - mu * mu) * a.v;
        }
        if (a.ms != 0.0) {
            double mb = m0[0] * a.bext[0] + m0[1] * a.bext[1] + m0[2] * a.bext[2];
            acc[2] -= a.ms * mb * a.v;
        }
        if (a.dmi_on && a.ms != 0.0) {
            double im[3], ip[3], jm[3], jp[3];
            auto load = [&](int64_t q, double *o) {
                o[0] = (double)mx[q]; o[1] = (double)my[q]; o[2] = (double)mz[q];
            };
            if (a.nx == 1) {
                for (int q = 0; q < 3; ++q) { im[q] = m0[q]; ip[q] = m0[q]; }
            } else {
                if (i == 0 || (GEOM && !geom[c - 1])) eg_x(m0, -1.0, a.eta, a.dx, im);
                else load(c - 1, im);
                if (i == a.nx - 1 || (GEOM && !geom[c + 1])) eg_x(m0, 1.0, a.eta, a.dx, ip);
                else load(c + 1, ip);
            }
            if (a.ny == 1) {
                for (int q = 0; q < 3; ++q) { jm[q] = m0[q]; jp[q] = m0[q]; }
            } else {
                if (j == 0 || (GEOM && !geom[c - a.nx])) eg_y(m0, -1.0, a.eta, a.dy, jm);
                else load(c - a.nx, jm);
                if (j == a.ny - 1 || (GEOM && !geom[c + a.nx])) eg_y(m0, 1.0, a.eta, a.dy, jp);
                else load(c + a.nx, jp);
            }
            double dmz_dx = a.nx > 1 ? (ip[2] - im[2]) / (2.0 * a.dx) : 0.0;
            double dmz_dy = a.ny > 1 ? (jp[2] - jm[2]) / (2.0 * a.dy) : 0.0;
            double dmx_dx = a.nx > 1 ? (ip[0] - im[0]) / (2.0 * a.dx) : 0.0;
            double dmy_dy = a.ny > 1 ? (jp[1] - jm[1]) / (2.0 * a.dy) : 0.0;
            double bx = a.pref * dmz_dx, by = a.pref * dmz_dy, bz = -a.pref * (dmx_dx + dmy_dy);
            double mb = m0[0] * bx + m0[1] * by + m0[2] * bz;
            acc[3] += -0.5 * a.ms * mb * a.v;
        }
        if (DEMAG && a.ms != 0.0) {
            double mb = m0[0] * (double)bd[c] + m0[1] * (double)bd[n + c] + m0[2] * (double)bd[2 * n + c];
            acc[4] += -0.5 * a.ms * mb * a.v;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        double v = wsum(acc[q]);
        if (lane == 0) red[q][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = red[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        partial[(int64_t)blockIdx.x * 5 + threadIdx.x] = s;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_torque(const T *__restrict__ m, const T *__restrict__ b, int64_t n, double *__restrict__ partial) {
    __shared__ double red[3][4];
    double mx = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double m0 = m[c], m1 = m[n + c], m2 = m[2 * n + c];
        const double b0 = b[c], b1 = b[n + c], b2 = b[2 * n + c];
        const double t0 = m1 * b2 - m2 * b1, t1 = m2 * b0 - m0 * b2, t2 = m0 * b1 - m1 * b0;
        const double q2 = t0 * t0 + t1 * t1 + t2 * t2;
        const double q = sqrt(q2);
        mx = fmax(mx, q);
        s1 += q;
        s2 += q2;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    mx = wmax(mx);
    s1 = wsum(s1);
    s2 = wsum(s2);
    if (lane == 0) { red[0][wid] = mx; red[1][wid] = s1; red[2][wid] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = red[0][0], r1 = red[1][0], r2 = red[2][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            r0 = fmax(r0, red[0][w]);
            r1 += red[1][w];
            r2 += red[2][w];
        }
        partial[(int64_t)blockIdx.x * 3] = r0;
        partial[(int64_t)blockIdx.x * 3 + 1] = r1;
        partial[(int64_t)blockIdx.x * 3 + 2] = r2;
    }
}

int red_blocks(int64_t n) {
    int64_t want = (n + 255) / 256;
    return (int)(want < kRelaxRedBlocks ? (want > 0 ? want : 1) : kRelaxRedBlocks);
}

}  // namespace

template <typename T>
void launch_rk23_err(const T *m, const T *k1, const T *k2, const T *k3, const T *k4, const T *m4,
                     const Rk23ErrCoef &c, int64_t n, double *partial, double *out, hipStream_t s) {
    const int nb = red_blocks(n);
    if (sizeof(T) == sizeof(float))
        hipLaunchKernelGGL((k_rk23_err<T, true>), dim3(nb), dim3(256), 0, s, m, k1, k2, k3, k4, m4, c, n, partial);
    else
        hipLaunchKernelGGL((k_rk23_err<T, false>), dim3(nb), dim3(256), 0, s, m, k1, k2, k3, k4, m4, c, n, partial);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rows_reduce, dim3(1), dim3(1024), 0, s, partial, nb, 1, 1, out);
    LLGMI_LAUNCH_CHECK();
}

template <typename T>
void launch_energy(const EnergyArgs &a, const T *m, const T *bd, const uint8_t *geom,
                   double *partial, double *out, hipStream_t s) {
    const int64_t n = (int64_t)a.nx * a.ny;
    const int nb = red_blocks(n);
#define LLGMI_EN(G, D) \
    hipLaunchKernelGGL((k_energy<T, G, D>), dim3(nb), dim3(256), 0, s, a, m, bd, geom, partial)
    if (geom) { if (bd) LLGMI_EN(true, true); else LLGMI_EN(true, false); }
    else { if (bd) LLGMI_EN(false, true); else LLGMI_EN(false, false); }
#undef LLGMI_EN
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rows_reduce, dim3(1), dim3(1024), 0, s, partial, nb, 5, 0, out);
    LLGMI_LAUNCH_CHECK();
}

template <typename T>
void launch_torque(const T *m, const T *b, int64_t n, double *partial, double *out, hipStream_t s) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL((k_torque<T>), dim3(nb), dim3(256), 0, s, m, b, n, partial);
    LLGMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rows_reduce, dim3(1), dim3(1024), 0, s, partial, nb, 3, 1, out);
    LLGMI_LAUNCH_CHECK();
}

#define LLGMI_INST(T)                                                                              \
    template void launch_rk23_err<T>(const T *, const T *, const T *, const T *, const T *,       \
                                     const T *, const Rk23ErrCoef &, int64_t, double *, double *, \
                                     hipStream_t);                                                 \
    template void launch_energy<T>(const EnergyArgs &, const T *, const T *, const uint8_t *,     \
                                   double *, double *, hipStream_t);                               \
    template void launch_torque<T>(const T *, const T *, int64_t, double *, double *, hipStream_t);
LLGMI_INST(float)
LLGMI_INST(double)
#undef LLGMI_INST

}  // namespace llgmi

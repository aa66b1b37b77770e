// FFT demag for the uniform 2D grid (Nz = 1), open boundaries.
//
// Same operator as llg-sim's Demag2D (effective_field/demag_fft_uniform.rs:434-769):
//   B_i = IFFT( K^_ij * FFT(Ms m_j) ) restricted to the physical region,
// with the MuMax-like face-charge kernel (demag_fft_uniform.rs:1013-1253)
// evaluated in FP64 on the GPU.
//
// MI355X layout (differs from the reference's six full complex 2D FFTs):
//  * Padded size px x py is the next power of two >= 2nx x 2ny.  A circular
//    convolution on any period >= 2n-1 equals the reference's linear
//    (2n-padded) convolution on the physical region.
//  * The kernels are even (xx, yy, zz) or odd-odd (xy) in real space, so their
//    transforms are real and quadrant-symmetric: only (px/2+1) x (py/2+1)
//    real 4-vectors are stored (xy changes sign for ky > py/2).
//  * Pass 1 (row R2C): two real rows per complex FFT, rows >= ny and
//    columns >= nx are zero and never read.  The half spectrum is written
//    column-major, S[c][kx][j], j < ny only, in packed form: px/2 complex
//    columns, column 0 holding Re X[0] + i Re X[px/2] (both are real for a
//    real row), so the column pass has exactly px/2 independent columns.
//  * Pass 2 (column): per kx, the three components are zero-extended to py,
//    transformed, multiplied by the 4-component tensor (scale 1/(px py)
//    folded in), inverse transformed; only j < ny is stored.  The packed
//    column 0 carries two real-data transforms at once; Kxy vanishes on both
//    kx = 0 and kx = px/2, so it needs only the diagonal entries of the two
//    kernel rows, combined through the Hermitian partner ky <-> py - ky.
//  * Pass 3 (row C2R): two Hermitian rows per complex inverse FFT; the first
//    nx outputs of each row are written as B_demag (SoA).
#include "common.hpp"
#include "fft_reg.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include <cstdlib>
#include "internal.hpp"
#include "local_cell.hpp"

#include <cmath>
#include <vector>

namespace llgmi {

// Block shapes, fixed per (precision, log2 length).  One sequence of length L
// is transformed by TPS = L/R0 threads holding R0 values each (fft_reg.hpp).
// Row passes hold PAIRS pairs of rows (one complex sequence per pair) in at
// most ~80 KB of LDS; the column pass holds the three components of one column.
__host__ __device__ constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__host__ __device__ constexpr int mini(int a, int b) { return a < b ? a : b; }

// Row blocks: as many row pairs as fit the LDS / thread budget (more pairs per
// block = longer contiguous runs in the column-major spectrum).
#ifndef LLGMI_ROW_LDS_KB
#define LLGMI_ROW_LDS_KB 80
#endif
#ifndef LLGMI_ROW_NT
#define LLGMI_ROW_NT 512
#endif
template <typename T, int LOG2L, int RQ = 0>
struct RowCfg {
    using P = Plan<LOG2L, RQ>;
    static constexpr int L = P::L, TPS = P::TPS, SLEN = lds_seq_len(L);
    static constexpr int PAIRS = clampi(mini((int)(LLGMI_ROW_LDS_KB * 1024 / (sizeof(cpx<T>) * SLEN)),
                                             LLGMI_ROW_NT / TPS), 1, 64);
    static constexpr int NT = PAIRS * TPS;
    static constexpr size_t LDS = sizeof(cpx<T>) * PAIRS * SLEN;
};

template <typename T, int LOG2L, int RQ = 0>
struct ColCfg {
    using P = Plan<LOG2L, RQ>;
    static constexpr int L = P::L, TPS = P::TPS, SLEN = lds_seq_len(L);
    static constexpr int NT = 2 * TPS;
    static constexpr size_t LDS = sizeof(cpx<T>) * 2 * SLEN;
};

// One-item-per-block column launches group G items (G * 2 sequences) per
// block when a sequence needs few threads, so that a block fills at least two
// waves (short columns would otherwise leave most lanes of a wave idle).
template <typename T, int LOG2L, int RQ>
struct ColGrp {
    using C = ColCfg<T, LOG2L, RQ>;
    static constexpr int G = 2 * C::TPS >= 128 ? 1 : 128 / (2 * C::TPS);
    static constexpr int NT = G * 2 * C::TPS;
    static constexpr size_t LDS = sizeof(cpx<T>) * 2 * G * C::SLEN;
};

template <int LOG2L>
struct ExtractCfg {
    using P = Plan<LOG2L>;
    static constexpr int L = P::L, TPS = P::TPS, SLEN = lds_seq_len(L);
    static constexpr int NT = TPS;
    static constexpr size_t LDS = sizeof(cpx<double>) * SLEN;
};

// ---------------------------------------------------------------------------
// Pass 1: row R2C.  Block = PAIRS pairs of rows of one component; two real rows
// form one complex sequence, loaded straight into registers.  Columns
// >= ncols_valid and rows >= nrows are zero and not read.
// dst[comp*dst_cstride + k*nrows + j] for k <= L/2, j < nrows.
// ---------------------------------------------------------------------------
// HALF: ncols_valid == L/2, so the upper half of every row is a compile-time zero.
// The block of component comp whose first row is j0.
template <typename T, typename Tin, int LOG2L, bool HALF, int RQ>
__device__ __forceinline__ void row_r2c_body(const Tin *__restrict__ src, int64_t src_cstride, int row_stride,
                                             int ncols_valid, int nrows, T scale, const cpx<T> *__restrict__ tw,
                                             cpx<T> *__restrict__ dst, int64_t dst_cstride, int comp, int j0) {
    using C = RowCfg<T, LOG2L, RQ>;
    using PL = typename C::P;
    constexpr int L = C::L, PAIRS = C::PAIRS, NT = C::NT, SLEN = C::SLEN, TPS = C::TPS;
    constexpr int NV = PL::NV;
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lds_raw);
    const int p = threadIdx.x / TPS, t = threadIdx.x - p * TPS;
    const int ja = j0 + 2 * p, jb = ja + 1;
    const Tin *ra = src + comp * src_cstride + (int64_t)ja * row_stride;
    const Tin *rb = ra + row_stride;
    cpx<T> v[NV];
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int i = PL::tindex(t, r);
        T re = T(0), im = T(0);
        if (HALF ? PL::lower(r) : i < ncols_valid) {
            if (ja < nrows) re = scale * T(ra[i]);
            if (jb < nrows) im = scale * T(rb[i]);
        }
        v[r] = {re, im};
    }
    cpx<T> *seq = buf + p * SLEN;
    plan_fft<T, false, LOG2L, RQ>(v, seq, t, tw);
    __syncthreads();
    {
        cpx<T> *wb = seq + lds_pad(t);
#pragma unroll
        for (int r = 0; r < NV; ++r) wb[lds_pad(r * TPS)] = v[r];
    }
    __syncthreads();
    constexpr int HXS = L / 2;  // packed columns
    constexpr int NR = 2 * PAIRS;
    cpx<T> *d = dst + comp * dst_cstride;
    for (int idx = threadIdx.x; idx < HXS * NR; idx += NT) {
        const int r = idx % NR, k = idx / NR;
        const int j = j0 + r;
        if (j >= nrows) continue;
        const int pp = r >> 1;
        cpx<T> X;
        if (k == 0) {
            // X_a[0] = Re Z[0], X_b[0] = Im Z[0]; X_a[L/2] = Re Z[L/2], X_b[L/2] = Im Z[L/2]
            const cpx<T> z0 = buf[pp * SLEN + lds_pad(0)];
            const cpx<T> zn = buf[pp * SLEN + lds_pad(L / 2)];
            X = (r & 1) == 0 ? cpx<T>{z0.x, zn.x} : cpx<T>{z0.y, zn.y};
        } else {
            const cpx<T> z = buf[pp * SLEN + lds_pad(k)];
            const cpx<T> zc = buf[pp * SLEN + lds_pad(L - k)];
            if ((r & 1) == 0) X = {T(0.5) * (z.x + zc.x), T(0.5) * (z.y - zc.y)};
            else X = {T(0.5) * (z.y + zc.y), T(0.5) * (zc.x - z.x)};
        }
        d[(int64_t)k * nrows + j] = X;
    }
}

template <typename T, typename Tin, int LOG2L, bool HALF, int RQ>
__global__ void __launch_bounds__((RowCfg<T, LOG2L, RQ>::NT))
k_row_r2c(const Tin *__restrict__ src, int64_t src_cstride, int row_stride, int ncols_valid,
          int nrows, T scale, const cpx<T> *__restrict__ tw, cpx<T> *__restrict__ dst,
          int64_t dst_cstride) {
    constexpr int PAIRS = RowCfg<T, LOG2L, RQ>::PAIRS;
    const int lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int comp = lb / gridDim.x;
    const int j0 = (lb - comp * gridDim.x) * PAIRS * 2;
    row_r2c_body<T, Tin, LOG2L, HALF, RQ>(src, src_cstride, row_stride, ncols_valid, nrows, scale, tw, dst,
                                          dst_cstride, comp, j0);
}

// ---------------------------------------------------------------------------
// Pass 2: column FFT, tensor multiply, inverse column FFT.
// A work item transforms two sequences, with TPS threads each:
//   items [0, hxs):            (x, y) of column kx = item   (coupled by Kxy)
//   items [hxs, hxs + hxs/2):  z of columns 2w' and 2w'+1   (Kzz only)
// Two sequences per block keep the block at 2*TPS threads and ~35 KB of LDS.
// The forward outputs stay in registers; x and y exchange through LDS for the
// off-diagonal Kxy terms.  PIPE = false: one item per block.  PIPE = true:
// persistent blocks (as many as are resident) step through the items, and
// the column data of a block's next item is loaded while the current one is
// transformed, so that loads, arithmetic and stores of different items overlap.
// ---------------------------------------------------------------------------
// One work item: v holds the thread's column values (zero where padded) on
// entry; the item's result is stored to col (null: nothing to store).
template <typename T, int LOG2L, bool HALF, int RQ>
__device__ __forceinline__ void col_item(cpx<T> *v, cpx<T> *col, int comp, int kx, int hxs, int ny, cpx<T> *buf,
                                         int sidx, int t, const cpx<T> *__restrict__ tw,
                                         const T *__restrict__ kern) {
    using C = ColCfg<T, LOG2L, RQ>;
    using PL = typename C::P;
    constexpr int L = C::L, SLEN = C::SLEN, TPS = C::TPS;
    constexpr int NV = PL::NV;
    constexpr int HY = L / 2 + 1;
    cpx<T> *seq = buf + sidx * SLEN;
    // kernel planes [Kxx, Kxy, Kyy, Kzz][hxs + 1][HY]: the (x, y) items read
    // two planes per sequence (diagonal + Kxy), the z items only Kzz; the
    // values are read before the forward transform so that their latency
    // overlaps it
    const int64_t plane = (int64_t)(hxs + 1) * HY;
    const T *pa, *pb;
    if (kx != 0) {
        pa = kern + (comp == 0 ? 0 : (comp == 1 ? 2 : 3)) * plane + (int64_t)kx * HY;
        pb = kern + plane + (int64_t)kx * HY;
    } else {
        pa = kern + (comp == 0 ? 0 : (comp == 1 ? 2 : 3)) * plane;
        pb = pa + (int64_t)hxs * HY;
    }
    const bool need_b = kx == 0 || comp != 2;
    T ka[NV], kb[NV];
    auto load_k = [&] {
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int ky = t + r * TPS;
            const int kq = ky < HY ? ky : L - ky;
            ka[r] = pa[kq];
            kb[r] = need_b ? pb[kq] : T(0);
        }
    };
    // FP32 power-of-two lengths: read before the transform; otherwise the
    // registers are kept for the data
    constexpr bool KPRE = sizeof(T) == 4 && !PL::F3;
    if constexpr (KPRE) load_k();
    plan_fft<T, false, LOG2L, RQ>(v, seq, t, tw);
    if constexpr (!KPRE) load_k();
    __syncthreads();
    {
        cpx<T> *wb = seq + lds_pad(t);
#pragma unroll
        for (int r = 0; r < NV; ++r) wb[lds_pad(r * TPS)] = v[r];
    }
    __syncthreads();
    if (kx != 0) {
        const cpx<T> *other = buf + (sidx ^ 1) * SLEN + lds_pad(t);
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int ky = t + r * TPS;
            const cpx<T> own = v[r];
            if (comp == 2) {
                const T kzz = ka[r];
                v[r] = {kzz * own.x, kzz * own.y};
            } else {
                const T kd = ka[r];
                const T kxy = ky < HY ? kb[r] : -kb[r];
                const cpx<T> oth = other[lds_pad(r * TPS)];
                if (comp == 0)
                    v[r] = {kd * own.x + kxy * oth.x, kd * own.y + kxy * oth.y};
                else
                    v[r] = {kxy * oth.x + kd * own.x, kxy * oth.y + kd * own.y};
            }
        }
    } else {
        // packed column: Z = A + iB with A, B the transforms of the real kx = 0
        // and kx = px/2 columns; Kxy vanishes on both, so per component
        // Y = Ka A + i Kb B = ((Ka+Kb)/2) Z + ((Ka-Kb)/2) conj(Z[-ky])
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int ky = t + r * TPS;
            const T a = ka[r], b = kb[r];
            const T c1 = T(0.5) * (a + b), c2 = T(0.5) * (a - b);
            const cpx<T> z = v[r];
            const cpx<T> zp = seq[lds_pad(ky == 0 ? 0 : L - ky)];
            v[r] = {c1 * z.x + c2 * zp.x, c1 * z.y - c2 * zp.y};
        }
    }
    plan_fft<T, true, LOG2L, RQ>(v, seq, t, tw);
    if (col) {
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int j = PL::tindex(t, r);
            if (HALF ? PL::lower(r) : j < ny) col[j] = v[r];
        }
    }
}

// Items wg * G .. wg * G + G - 1 of the column pass (items [0, nwork)), one
// block (ColGrp); each item has its own two sequences of LDS.
template <typename T, int LOG2L, bool HALF, int RQ>
__device__ __forceinline__ void col_work(cpx<T> *__restrict__ S, int hxs, int ny, const cpx<T> *__restrict__ tw,
                                         const T *__restrict__ kern, int nwork, int wg) {
    using C = ColCfg<T, LOG2L, RQ>;
    using PL = typename C::P;
    constexpr int TPS = C::TPS, SLEN = C::SLEN;
    constexpr int NV = PL::NV;
    constexpr int G = ColGrp<T, LOG2L, RQ>::G;
    extern __shared__ unsigned char lds_raw[];
    const int g = threadIdx.x / (2 * TPS), tl = threadIdx.x - g * 2 * TPS;
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lds_raw) + g * 2 * SLEN;
    const int sidx = tl / TPS, t = tl - sidx * TPS;
    const int w = wg * G + g;
    const bool xy = w < hxs;
    const int comp = xy ? sidx : 2;
    int kx = xy ? w : 2 * (w - hxs) + sidx;
    const bool ok = w < nwork && kx < hxs;
    if (!ok) kx = 0;
    cpx<T> *col = ok ? S + ((int64_t)comp * hxs + kx) * ny : nullptr;
    cpx<T> v[NV];
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int j = PL::tindex(t, r);
        v[r] = (col && (HALF ? PL::lower(r) : j < ny)) ? col[j] : cpx<T>{T(0), T(0)};
    }
    col_item<T, LOG2L, HALF, RQ>(v, col, comp, kx, hxs, ny, buf, sidx, t, tw, kern);
}

template <typename T, int LOG2L, bool HALF, int RQ, bool PIPE>
__global__ void __launch_bounds__((PIPE ? ColCfg<T, LOG2L, RQ>::NT : ColGrp<T, LOG2L, RQ>::NT))
k_col_demag(cpx<T> *__restrict__ S, int hxs, int ny, const cpx<T> *__restrict__ tw,
            const T *__restrict__ kern, int nwork) {
    using C = ColCfg<T, LOG2L, RQ>;
    constexpr int TPS = C::TPS;
    constexpr int R0 = C::P::R0;
    constexpr int NL = HALF ? R0 / 2 : R0;   // values a thread loads per sequence
    static_assert(!PIPE || !C::P::F3, "the persistent column pass is built for power-of-two lengths");
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lds_raw);
    const int sidx = threadIdx.x / TPS, t = threadIdx.x - sidx * TPS;
    // column of this thread's sequence in item w (null: none); comp and kx
    // as the item defines them (kx clamped to a valid column when there is none)
    auto column = [&](int w, int &comp, int &kx) -> cpx<T> * {
        const bool xy = w < hxs;
        comp = xy ? sidx : 2;
        kx = xy ? w : 2 * (w - hxs) + sidx;
        const bool ok = w < nwork && kx < hxs;
        if (!ok) kx = 0;
        return ok ? S + ((int64_t)comp * hxs + kx) * ny : nullptr;
    };
    if constexpr (!PIPE) {
        col_work<T, LOG2L, HALF, RQ>(S, hxs, ny, tw, kern, nwork, (int)blockIdx.x);
    } else {
        // persistent: the next item's column data is in flight while the
        // current item is transformed
        int w = blockIdx.x;
        cpx<T> nxt[NL];
        {
            int c0, k0;
            const cpx<T> *c = column(w, c0, k0);
#pragma unroll
            for (int r = 0; r < NL; ++r) {
                const int j = t + r * TPS;
                nxt[r] = (c && (HALF || j < ny)) ? c[j] : cpx<T>{T(0), T(0)};
            }
        }
        for (; w < nwork; w += (int)gridDim.x) {
            int comp, kx;
            cpx<T> *col = column(w, comp, kx);
            cpx<T> v[R0];
#pragma unroll
            for (int r = 0; r < R0; ++r) v[r] = r < NL ? nxt[r] : cpx<T>{T(0), T(0)};
            {
                int cn, kn;
                const cpx<T> *c = column(w + (int)gridDim.x, cn, kn);
#pragma unroll
                for (int r = 0; r < NL; ++r) {
                    const int j = t + r * TPS;
                    nxt[r] = (c && (HALF || j < ny)) ? c[j] : cpx<T>{T(0), T(0)};
                }
            }
            col_item<T, LOG2L, HALF, RQ>(v, col, comp, kx, hxs, ny, buf, sidx, t, tw, kern);
        }
    }
}

// ---------------------------------------------------------------------------
// Split column pass, for columns whose item (two padded sequences) does not
// fit in a CU's LDS (FP64 columns of 2^13): one sequence per block and an HBM
// round trip between the transforms.  k_colsplit_fwd transforms one column
// (zero past ny) and stores its full-length spectrum in F[comp][kx][L];
// k_colsplit_mul_inv forms the product with the kernel from F (its own column
// and, for the x and y components, the other one at the same ky), transforms
// it back and stores the first ny values into S.  The products are those of
// col_item.
// ---------------------------------------------------------------------------
template <typename T, int LOG2L, bool HALF, int RQ>
__global__ void __launch_bounds__((RegPlan<LOG2L, RQ>::TPS))
k_colsplit_fwd(const cpx<T> *__restrict__ S, cpx<T> *__restrict__ F, int ny, const cpx<T> *__restrict__ tw) {
    static_assert(!Plan<LOG2L, RQ>::F3, "the split column pass is built for power-of-two lengths");
    using P = Plan<LOG2L, RQ>;
    constexpr int L = P::L, TPS = P::TPS, R0 = P::R0;
    constexpr int NL = HALF ? R0 / 2 : R0;
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *seq = reinterpret_cast<cpx<T> *>(lds_raw);
    const int t = threadIdx.x;
    const int64_t c = blockIdx.x;  // comp * hxs + kx
    const cpx<T> *col = S + c * ny;
    cpx<T> v[R0];
#pragma unroll
    for (int r = 0; r < R0; ++r) {
        const int j = t + r * TPS;
        v[r] = (r < NL && (HALF || j < ny)) ? col[j] : cpx<T>{T(0), T(0)};
    }
    reg_fft<T, false, LOG2L, RQ>(v, seq, t, tw);
    cpx<T> *f = F + c * L;
#pragma unroll
    for (int r = 0; r < R0; ++r) f[t + r * TPS] = v[r];
}

template <typename T, int LOG2L, bool HALF, int RQ>
__global__ void __launch_bounds__((RegPlan<LOG2L, RQ>::TPS))
k_colsplit_mul_inv(cpx<T> *__restrict__ S, const cpx<T> *__restrict__ F, int hxs, int ny,
                   const cpx<T> *__restrict__ tw, const T *__restrict__ kern) {
    static_assert(!Plan<LOG2L, RQ>::F3, "the split column pass is built for power-of-two lengths");
    using P = Plan<LOG2L, RQ>;
    constexpr int L = P::L, TPS = P::TPS, R0 = P::R0;
    constexpr int HY = L / 2 + 1;
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *seq = reinterpret_cast<cpx<T> *>(lds_raw);
    const int t = threadIdx.x;
    const int comp = (int)blockIdx.x / hxs, kx = (int)blockIdx.x - comp * hxs;
    const cpx<T> *own = F + ((int64_t)comp * hxs + kx) * L;
    // kernel planes [Kxx, Kxy, Kyy, Kzz][hxs + 1][HY], as in col_item
    const int64_t plane = (int64_t)(hxs + 1) * HY;
    const int dplane = comp == 0 ? 0 : (comp == 1 ? 2 : 3);
    cpx<T> v[R0];
    if (kx != 0) {
        const T *pa = kern + dplane * plane + (int64_t)kx * HY;
        const T *pb = kern + plane + (int64_t)kx * HY;
        const cpx<T> *oth = comp < 2 ? F + ((int64_t)(1 - comp) * hxs + kx) * L : own;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int ky = t + r * TPS;
            const int kq = ky < HY ? ky : L - ky;
            const cpx<T> o = own[ky];
            if (comp == 2) {
                const T kzz = pa[kq];
                v[r] = {kzz * o.x, kzz * o.y};
            } else {
                const T kd = pa[kq];
                const T kxy = ky < HY ? pb[kq] : -pb[kq];
                const cpx<T> q = oth[ky];
                if (comp == 0)
                    v[r] = {kd * o.x + kxy * q.x, kd * o.y + kxy * q.y};
                else
                    v[r] = {kxy * q.x + kd * o.x, kxy * q.y + kd * o.y};
            }
        }
    } else {
        // packed DC / Nyquist column, as in col_item
        const T *pa = kern + dplane * plane;
        const T *pb = pa + (int64_t)hxs * HY;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int ky = t + r * TPS;
            const int kq = ky < HY ? ky : L - ky;
            const T a = pa[kq], b = pb[kq];
            const T c1 = T(0.5) * (a + b), c2 = T(0.5) * (a - b);
            const cpx<T> z = own[ky];
            const cpx<T> zp = own[(L - ky) & (L - 1)];
            v[r] = {c1 * z.x + c2 * zp.x, c1 * z.y - c2 * zp.y};
        }
    }
    reg_fft<T, true, LOG2L, RQ>(v, seq, t, tw);
    cpx<T> *col = S + ((int64_t)comp * hxs + kx) * ny;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
        const int j = t + r * TPS;
        if (HALF ? r < R0 / 2 : j < ny) col[j] = v[r];
    }
}

// ---------------------------------------------------------------------------
// Pass 3: row C2R.  Block = PAIRS pairs of rows of one component.  The k = 0
// and k = L/2 bins use their real parts only, which is what taking Re() of the
// reference's full complex inverse transform amounts to.
// ---------------------------------------------------------------------------
// HALF: nx == L/2, so only the lower half of each inverse-transformed row is
// formed and stored.
template <typename T, int LOG2L, bool HALF, int RQ>
__device__ __forceinline__ void row_c2r_body(const cpx<T> *__restrict__ S, int hxs, int nx, int ny,
                                             const cpx<T> *__restrict__ tw, T *__restrict__ out, int64_t n,
                                             int comp, int j0) {
    using C = RowCfg<T, LOG2L, RQ>;
    using PL = typename C::P;
    constexpr int L = C::L, PAIRS = C::PAIRS, NT = C::NT, SLEN = C::SLEN, TPS = C::TPS;
    constexpr int NV = PL::NV;
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lds_raw);
    const cpx<T> *s = S + (int64_t)comp * hxs * ny;
    constexpr int NYQ = L >> 1;
    for (int idx = threadIdx.x; idx < NYQ * PAIRS; idx += NT) {
        const int p = idx % PAIRS, k = idx / PAIRS;
        const int ja = j0 + 2 * p, jb = ja + 1;
        const cpx<T> xa = ja < ny ? s[(int64_t)k * ny + ja] : cpx<T>{T(0), T(0)};
        const cpx<T> xb = jb < ny ? s[(int64_t)k * ny + jb] : cpx<T>{T(0), T(0)};
        cpx<T> *row = buf + p * SLEN;
        if (k == 0) {
            // packed: xa = (Re X_a[0], Re X_a[L/2]); the imaginary parts of the
            // two real bins are dropped, as Re() of a full inverse transform does
            row[lds_pad(0)] = {xa.x, xb.x};
            row[lds_pad(NYQ)] = {xa.y, xb.y};
        } else {
            row[lds_pad(k)] = {xa.x - xb.y, xa.y + xb.x};
            row[lds_pad(L - k)] = {xa.x + xb.y, xb.x - xa.y};
        }
    }
    __syncthreads();
    const int p = threadIdx.x / TPS, t = threadIdx.x - p * TPS;
    cpx<T> *seq = buf + p * SLEN;
    cpx<T> v[NV];
    {
        const cpx<T> *rb = seq + lds_pad(t);
#pragma unroll
        for (int r = 0; r < NV; ++r) v[r] = rb[lds_pad(r * TPS)];
    }
    plan_fft<T, true, LOG2L, RQ>(v, seq, t, tw);
    const int ja = j0 + 2 * p;
    T *o = out + comp * n;
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int i = PL::tindex(t, r);
        if (HALF ? PL::lower(r) : i < nx) {
            if (ja < ny) o[(int64_t)ja * nx + i] = v[r].x;
            if (ja + 1 < ny) o[(int64_t)(ja + 1) * nx + i] = v[r].y;
        }
    }
}

template <typename T, int LOG2L, bool HALF, int RQ>
__global__ void __launch_bounds__((RowCfg<T, LOG2L, RQ>::NT))
k_row_c2r(const cpx<T> *__restrict__ S, int hxs, int nx, int ny, const cpx<T> *__restrict__ tw,
          T *__restrict__ out, int64_t n) {
    constexpr int PAIRS = RowCfg<T, LOG2L, RQ>::PAIRS;
    const int lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int comp = lb / gridDim.x;
    const int j0 = (lb - comp * gridDim.x) * PAIRS * 2;
    row_c2r_body<T, LOG2L, HALF, RQ>(S, hxs, nx, ny, tw, out, n, comp, j0);
}

// ---------------------------------------------------------------------------
// Pass 3 fused with the local field: a block takes PAIRS pairs of rows for the
// three components, inverse-transforms them, parks the real B_demag rows in
// LDS and evaluates the local field of those rows (Zeeman, exchange,
// anisotropy, DMI, + B_demag, optional torque) with the same per-cell code as
// the stand-alone local-field kernel, so B_demag never goes through HBM.
// ---------------------------------------------------------------------------
template <typename T, int LOG2L, int RQ>
struct FusedCfg {
    using P = Plan<LOG2L, RQ>;
    static constexpr int L = P::L, TPS = P::TPS, SLEN = lds_seq_len(L);
    static constexpr size_t kLdsCU = 160 * 1024;
    static constexpr int PAIRS = (3 * 2 * TPS <= 1024 && sizeof(cpx<T>) * 3 * 2 * SLEN <= kLdsCU) ? 2 : 1;
    static constexpr int NT = 3 * PAIRS * TPS;
    static constexpr size_t LDS = sizeof(cpx<T>) * 3 * PAIRS * SLEN;
    static constexpr bool OK = NT <= 1024 && LDS <= kLdsCU;
};

template <typename T, int LG, int RQ>
constexpr bool fused_ok() { return FusedCfg<T, LG, RQ>::OK; }

// The batched fused pass groups as many row pairs per block as keep it at
// 3 * PAIRS * TPS = 768 threads (within the LDS and thread limits), so that
// films of different lengths share one launch.  Rows are transformed and
// their cells evaluated the same way for any PAIRS.
template <typename T, int LOG2L, int RQ>
struct BatchFusedCfg {
    using P = Plan<LOG2L, RQ>;
    static constexpr int L = P::L, TPS = P::TPS, SLEN = lds_seq_len(L);
    static constexpr size_t kLdsCU = 160 * 1024;
    static constexpr int fit(int p) {
        return (p > 1 && (3 * p * TPS > 1024 || sizeof(cpx<T>) * 3 * p * SLEN > kLdsCU)) ? fit(p / 2) : p;
    }
    static constexpr int PAIRS = fit(TPS >= 256 ? 1 : 256 / TPS);
    static constexpr int NT = 3 * PAIRS * TPS;
    static constexpr size_t LDS = sizeof(cpx<T>) * 3 * PAIRS * SLEN;
    static constexpr bool OK = NT <= 1024 && LDS <= kLdsCU;
};

template <typename T, int LOG2L, bool HALF, int RQ, bool GEOM, bool TORQUE, typename C = FusedCfg<T, LOG2L, RQ>>
__device__ __forceinline__ void row_c2r_local_body(const cpx<T> *__restrict__ S, int hxs, int ny,
                                                   const cpx<T> *__restrict__ tw, const LocalArgs<T> &a,
                                                   const T *__restrict__ m, const uint8_t *__restrict__ geom,
                                                   T *__restrict__ b, T *__restrict__ kout, int j0) {
    using PL = typename C::P;
    constexpr int L = C::L, PAIRS = C::PAIRS, NT = C::NT, SLEN = C::SLEN, TPS = C::TPS;
    constexpr int NV = PL::NV;
    constexpr int NYQ = L >> 1;
    extern __shared__ unsigned char lds_raw[];
    cpx<T> *buf = reinterpret_cast<cpx<T> *>(lds_raw);
    const int nx = a.nx;
    for (int idx = threadIdx.x; idx < NYQ * 3 * PAIRS; idx += NT) {
        const int p = idx % PAIRS, rem = idx / PAIRS;
        const int comp = rem % 3, k = rem / 3;
        const cpx<T> *s = S + (int64_t)comp * hxs * ny;
        const int ja = j0 + 2 * p, jb = ja + 1;
        const cpx<T> xa = ja < ny ? s[(int64_t)k * ny + ja] : cpx<T>{T(0), T(0)};
        const cpx<T> xb = jb < ny ? s[(int64_t)k * ny + jb] : cpx<T>{T(0), T(0)};
        cpx<T> *row = buf + (comp * PAIRS + p) * SLEN;
        if (k == 0) {
            row[lds_pad(0)] = {xa.x, xb.x};
            row[lds_pad(NYQ)] = {xa.y, xb.y};
        } else {
            row[lds_pad(k)] = {xa.x - xb.y, xa.y + xb.x};
            row[lds_pad(L - k)] = {xa.x + xb.y, xb.x - xa.y};
        }
    }
    __syncthreads();
    const int sq = threadIdx.x / TPS, t = threadIdx.x - sq * TPS;
    cpx<T> *seq = buf + sq * SLEN;
    cpx<T> v[NV];
    {
        const cpx<T> *rb = seq + lds_pad(t);
#pragma unroll
        for (int r = 0; r < NV; ++r) v[r] = rb[lds_pad(r * TPS)];
    }
    plan_fft<T, true, LOG2L, RQ>(v, seq, t, tw);
    __syncthreads();   // every sequence is out of LDS before the rows go in
    T *bdl = reinterpret_cast<T *>(lds_raw);   // [3][2 PAIRS][NYQ]
    {
        const int comp = sq / PAIRS, p = sq - comp * PAIRS;
        T *ra = bdl + (comp * 2 * PAIRS + 2 * p) * NYQ;
        T *rb = ra + NYQ;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            const int i = PL::tindex(t, r);
            if (HALF ? PL::lower(r) : i < nx) {
                ra[i] = v[r].x;
                rb[i] = v[r].y;
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * PAIRS * nx; q += NT) {
        const int rr = q / nx, i = q - rr * nx;
        const int j = j0 + rr;
        if (j >= ny) continue;
        const T bd3[3] = {bdl[(0 * 2 * PAIRS + rr) * NYQ + i], bdl[(1 * 2 * PAIRS + rr) * NYQ + i],
                          bdl[(2 * 2 * PAIRS + rr) * NYQ + i]};
        local_cell<T, GEOM, true, TORQUE>(a, m, geom, i, j, bd3, b, kout);
    }
}

template <typename T, int LOG2L, bool HALF, int RQ, bool GEOM, bool TORQUE>
__global__ void __launch_bounds__((FusedCfg<T, LOG2L, RQ>::NT))
k_row_c2r_local(const cpx<T> *__restrict__ S, int hxs, int ny, const cpx<T> *__restrict__ tw, LocalArgs<T> a,
                const T *__restrict__ m, const uint8_t *__restrict__ geom, T *__restrict__ b,
                T *__restrict__ kout) {
    constexpr int PAIRS = FusedCfg<T, LOG2L, RQ>::PAIRS;
    const int j0 = xcd_remap(blockIdx.x, gridDim.x) * PAIRS * 2;
    row_c2r_local_body<T, LOG2L, HALF, RQ, GEOM, TORQUE>(S, hxs, ny, tw, a, m, geom, b, kout, j0);
}

// ---------------------------------------------------------------------------
// One-time k-space kernel: forward column FFT (FP64) of the row spectra of the
// real-space kernels and extraction of the real quadrant, times the inverse
// transform's 1 / (px py) (a power of two, so the products of the column pass
// are exactly what a separate scaling would give).  Block = (kx, comp).
// ---------------------------------------------------------------------------
template <typename Tout, int LOG2L>
__global__ void __launch_bounds__((ExtractCfg<LOG2L>::NT))
k_col_extract(const cpx<double> *__restrict__ Sk, int hxs, const cpx<double> *__restrict__ tw,
              Tout *__restrict__ kern, double scale) {
    using C = ExtractCfg<LOG2L>;
    using PL = typename C::P;
    constexpr int L = C::L, TPS = C::TPS;
    constexpr int NV = PL::NV;
    extern __shared__ unsigned char lds_raw[];
    cpx<double> *buf = reinterpret_cast<cpx<double> *>(lds_raw);
    const int kx = blockIdx.x, comp = blockIdx.y;
    const int t = threadIdx.x;
    const cpx<double> *s = Sk + ((int64_t)comp * hxs + kx) * L;
    cpx<double> v[NV];
#pragma unroll
    for (int r = 0; r < NV; ++r) v[r] = s[PL::tindex(t, r)];
    plan_fft<double, false, LOG2L>(v, buf, t, tw);
    constexpr int HY = L / 2 + 1;
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const int ky = t + r * TPS;
        if (ky < HY) {
            const int64_t plane = (int64_t)(hxs + 1) * HY;
            kern[comp * plane + (int64_t)kx * HY + ky] = Tout(v[r].x * scale);
            // packed column 0: the imaginary part is the transform of the kx = px/2
            // column (both transforms are real for the even/odd-odd kernels)
            if (kx == 0) kern[comp * plane + (int64_t)hxs * HY + ky] = Tout(v[r].y * scale);
        }
    }
}

// ---------------------------------------------------------------------------
// MuMax-like face-charge quadrature (demag_fft_uniform.rs:1139-1253), FP64.
// ---------------------------------------------------------------------------
struct QuadSetup {
    int64_t nv, nw, nx, ny, nz;
    double charge, pu1, cell[3];
    int u, v, w;
};

__host__ __device__ inline double delta_cell_d(int64_t d) {
    double a = (double)(d < 0 ? -d : d);
    if (a > 0.0) a -= 1.0;
    return a;
}
// (displacement, axis) sums longer than this go to k_quad_heavy
static constexpr int64_t kQuadHeavy = int64_t(1) << 16;
static constexpr int kQuadChunk = 256;
static constexpr int kQuadHeavyScan = 256;   // heavy sums are looked for in [0, 256)^2

__host__ __device__ inline int64_t mumax_n_d(double x) {
    return (int64_t)floor(fmax(x, 1.0) + 0.5);
}

__host__ __device__ inline QuadSetup quad_setup(int u, double dx, double dy, double dz,
                                                int64_t sx, int64_t sy, double accuracy) {
#pragma clang fp contract(off)
    QuadSetup q;
    q.u = u; q.v = (u + 1) % 3; q.w = (u + 2) % 3;
    q.cell[0] = dx; q.cell[1] = dy; q.cell[2] = dz;
    double lmin = fmin(fmin(dx, dy), dz);
    double dxm = delta_cell_d(sx) * dx, dym = delta_cell_d(sy) * dy, dzm = 0.0 * dz;
    double d = sqrt(dxm * dxm + dym * dym + dzm * dzm);
    if (d == 0.0) d = lmin;
    double max_size = d / accuracy;
    q.nx = mumax_n_d(dx / max_size);
    q.ny = mumax_n_d(dy / max_size);
    q.nz = mumax_n_d(dz / max_size);
    q.nv = 2 * mumax_n_d(q.cell[q.v] / max_size);
    q.nw = 2 * mumax_n_d(q.cell[q.w] / max_size);
    double scale = 1.0 / (double)(q.nv * q.nw * q.nx * q.ny * q.nz);
    q.charge = q.cell[q.v] * q.cell[q.w] * scale;
    q.pu1 = 0.5 * q.cell[u];
    return q;
}

__host__ __device__ inline int64_t quad_points(double dx, double dy, double dz, int64_t sx,
                                               int64_t sy, double accuracy) {
    int64_t t = 0;
    for (int u = 0; u < 3; ++u) {
        QuadSetup q = quad_setup(u, dx, dy, dz, sx, sy, accuracy);
        t += q.nv * q.nw * q.nx * q.ny * q.nz;
    }
    return t;
}

// One sample point's contribution (hx, hy, hz) for the pole pair.
__device__ __forceinline__ void quad_point(const QuadSetup &q, double rx, double ry, double rz,
                                           const double pole_vw[3], double out[3]) {
#pragma clang fp contract(off)
    const double PI_D = 3.14159265358979323846;
    double pole[3] = {pole_vw[0], pole_vw[1], pole_vw[2]};
    pole[q.u] = q.pu1;
    double r1x = rx - pole[0], r1y = ry - pole[1], r1z = rz - pole[2];
    double r1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
    double q1 = q.charge / (4.0 * PI_D * r1 * r1 * r1);
    double hx1 = r1x * q1, hy1 = r1y * q1, hz1 = r1z * q1;
    pole[q.u] = -q.pu1;
    double r2x = rx - pole[0], r2y = ry - pole[1], r2z = rz - pole[2];
    double r2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
    double q2 = -q.charge / (4.0 * PI_D * r2 * r2 * r2);
    out[0] = hx1 + r2x * q2;
    out[1] = hy1 + r2y * q2;
    out[2] = hz1 + r2z * q2;
}

// Sum over sample points (ax, ay, az) for the pole pair at (pv, pw), added to h
// in the reference's order.  The contributions of QB consecutive z samples are
// formed first (independent, so their square roots and divisions overlap in
// one thread) and then added one by one: same values, same order.
__device__ inline void quad_inner(const QuadSetup &q, const double rc[3], double pv, double pw,
                                  double h[3]) {
#pragma clang fp contract(off)
    constexpr int QB = 4;
    double pole[3] = {0.0, 0.0, 0.0};
    pole[q.v] = pv;
    pole[q.w] = pw;
    for (int64_t ax = 0; ax < q.nx; ++ax) {
        double rx = rc[0] - 0.5 * q.cell[0] + q.cell[0] / (2.0 * (double)q.nx) + (double)ax * (q.cell[0] / (double)q.nx);
        for (int64_t ay = 0; ay < q.ny; ++ay) {
            double ry = rc[1] - 0.5 * q.cell[1] + q.cell[1] / (2.0 * (double)q.ny) + (double)ay * (q.cell[1] / (double)q.ny);
            int64_t az0 = 0;
            for (; az0 + QB <= q.nz; az0 += QB) {   // full chunks: straight-line code
                double c[QB][3];
#pragma unroll
                for (int t = 0; t < QB; ++t) {
                    const int64_t az = az0 + t;
                    double rz = rc[2] - 0.5 * q.cell[2] + q.cell[2] / (2.0 * (double)q.nz) + (double)az * (q.cell[2] / (double)q.nz);
                    quad_point(q, rx, ry, rz, pole, c[t]);
                }
#pragma unroll
                for (int t = 0; t < QB; ++t) {
                    h[0] += c[t][0];
                    h[1] += c[t][1];
                    h[2] += c[t][2];
                }
            }
            for (; az0 < q.nz; ++az0) {
                double c[3];
                double rz = rc[2] - 0.5 * q.cell[2] + q.cell[2] / (2.0 * (double)q.nz) + (double)az0 * (q.cell[2] / (double)q.nz);
                quad_point(q, rx, ry, rz, pole, c);
                h[0] += c[0];
                h[1] += c[1];
                h[2] += c[2];
            }
        }
    }
}

__device__ inline double pole_pos(double c, int64_t n, int64_t i) {
#pragma clang fp contract(off)
    return -0.5 * c + c / (2.0 * (double)n) + (double)i * (c / (double)n);
}

// One thread per (base-quadrant displacement, source axis), summing serially in
// the reference's loop order with FP contraction off, so every real-space
// kernel value is the reference's own rounding of the face-charge sum (these
// sums cancel heavily near the origin; a reordered sum differs at ~1e-12).
// Threads are ordered by displacement, so the few heavy near-field sums share
// the first waves.
__global__ void __launch_bounds__(256)
k_quad(double *__restrict__ base, int nx, int ny, double dx, double dy, double dz,
       double accuracy, double mu0) {
#pragma clang fp contract(off)
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)nx * ny * 3) return;
    int64_t t = g / 3;
    int u = (int)(g - 3 * t);
    int64_t sx = t % nx, sy = t / nx;
    double rc[3] = {(double)sx * dx, (double)sy * dy, 0.0};
    QuadSetup q = quad_setup(u, dx, dy, dz, sx, sy, accuracy);
    if (sx < kQuadHeavyScan && sy < kQuadHeavyScan && q.nv * q.nw * q.nx * q.ny * q.nz > kQuadHeavy)
        return;   // k_quad_heavy takes it
    double h[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < q.nv; ++i) {
        double pv = pole_pos(q.cell[q.v], q.nv, i);
        for (int64_t j = 0; j < q.nw; ++j)
            quad_inner(q, rc, pv, pole_pos(q.cell[q.w], q.nw, j), h);
    }
    if (u == 0) base[t * 4 + 0] = mu0 * h[0];
    else if (u == 1) { base[t * 4 + 1] = mu0 * h[0]; base[t * 4 + 2] = mu0 * h[1]; }
    else base[t * 4 + 3] = mu0 * h[2];
}

// The near-field displacements whose sums exceed kQuadHeavy points: one block
// per (displacement, axis).  Its threads evaluate kQuadChunk consecutive points
// of the reference's loop order (pole v, pole w, x, y, z; z fastest) at once
// and one thread adds them to the sum in that order, so the sum is the serial
// one, bit for bit, while the square roots and divisions run 256 wide.
__global__ void __launch_bounds__(256)
k_quad_heavy(double *__restrict__ base, const int *__restrict__ items, int nx_base, double dx, double dy,
             double dz, double accuracy, double mu0) {
#pragma clang fp contract(off)
    __shared__ double c[3][kQuadChunk];
    const int item = items[blockIdx.x];
    const int u = item % 3;
    const int64_t t = item / 3;
    const int64_t sx = t % nx_base, sy = t / nx_base;
    const double rc[3] = {(double)sx * dx, (double)sy * dy, 0.0};
    const QuadSetup q = quad_setup(u, dx, dy, dz, sx, sy, accuracy);
    const int64_t total = q.nv * q.nw * q.nx * q.ny * q.nz;
    double h0 = 0.0, h1 = 0.0, h2 = 0.0;
    for (int64_t c0 = 0; c0 < total; c0 += kQuadChunk) {
        const int64_t f = c0 + threadIdx.x;
        if (f < total) {
            int64_t r = f;
            const int64_t az = r % q.nz; r /= q.nz;
            const int64_t ay = r % q.ny; r /= q.ny;
            const int64_t ax = r % q.nx; r /= q.nx;
            const int64_t j = r % q.nw, i = r / q.nw;
            double pole[3] = {0.0, 0.0, 0.0};
            pole[q.v] = pole_pos(q.cell[q.v], q.nv, i);
            pole[q.w] = pole_pos(q.cell[q.w], q.nw, j);
            const double rx = rc[0] - 0.5 * q.cell[0] + q.cell[0] / (2.0 * (double)q.nx) + (double)ax * (q.cell[0] / (double)q.nx);
            const double ry = rc[1] - 0.5 * q.cell[1] + q.cell[1] / (2.0 * (double)q.ny) + (double)ay * (q.cell[1] / (double)q.ny);
            const double rz = rc[2] - 0.5 * q.cell[2] + q.cell[2] / (2.0 * (double)q.nz) + (double)az * (q.cell[2] / (double)q.nz);
            double o[3];
            quad_point(q, rx, ry, rz, pole, o);
            c[0][threadIdx.x] = o[0];
            c[1][threadIdx.x] = o[1];
            c[2][threadIdx.x] = o[2];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = (int)(total - c0 < kQuadChunk ? total - c0 : kQuadChunk);
            for (int e = 0; e < cnt; ++e) {
                h0 += c[0][e];
                h1 += c[1][e];
                h2 += c[2][e];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (u == 0) base[t * 4 + 0] = mu0 * h0;
        else if (u == 1) { base[t * 4 + 1] = mu0 * h0; base[t * 4 + 2] = mu0 * h1; }
        else base[t * 4 + 3] = mu0 * h2;
    }
}

// Parity reflection into the padded real-space arrays (demag_fft_uniform.rs:1029-1071):
// xx, yy, zz even; xy odd-odd and zero on both axes; bins outside |s| <= n-1 zero.
__global__ void k_reflect(const double *__restrict__ base, int nx, int ny, int px, int py,
                          double *__restrict__ kr) {
    int ix = blockIdx.x * blockDim.x + threadIdx.x;
    int iy = blockIdx.y;
    if (ix >= px) return;
    int sx, sy;
    bool okx = true, oky = true;
    if (ix < nx) sx = ix;
    else if (ix >= px - nx + 1) sx = ix - px;
    else { sx = 0; okx = false; }
    if (iy < ny) sy = iy;
    else if (iy >= py - ny + 1) sy = iy - py;
    else { sy = 0; oky = false; }
    int64_t np = (int64_t)px * py, o = (int64_t)iy * px + ix;
    if (!okx || !oky) {
        for (int c = 0; c < 4; ++c) kr[c * np + o] = 0.0;
        return;
    }
    int ax = sx < 0 ? -sx : sx, ay = sy < 0 ? -sy : sy;
    const double *k = base + ((int64_t)ay * nx + ax) * 4;
    double sg = (sx == 0 || sy == 0) ? 0.0 : (double)((sx > 0) == (sy > 0) ? 1 : -1);
    kr[0 * np + o] = k[0];
    kr[1 * np + o] = sg * k[1];
    kr[2 * np + o] = k[2];
    kr[3 * np + o] = k[3];
}

// Periodic images (build_kernel_realspace_pbc, demag_fft_uniform.rs:941-1005):
// base(sx, sy) = sum of the kernel over every displacement that folds onto bin
// (sx, sy) -- sx + k nx with |sx + k nx| <= ex - 1 on a periodic axis (ex =
// nx pbc_x), sx alone on an open one -- summed y-major in ascending order as the
// reference accumulates its bins.  ext holds the kernel on the non-negative
// quadrant [0, ex) x [0, ey); a negative displacement takes the parity of each
// component (xx, yy, zz even; xy odd in x and in y).  The folded bins keep that
// parity, so the open-boundary reflection below applies unchanged.
__global__ void k_fold_images(const double *__restrict__ ext, int ex, int ey, int nx, int ny, int pbx,
                              int pby, double *__restrict__ base) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int sx = (int)(t % nx), sy = (int)(t / nx);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int kyl = pby > 0 ? -pby : 0, kyh = pby > 0 ? pby : 0;
    const int kxl = pbx > 0 ? -pbx : 0, kxh = pbx > 0 ? pbx : 0;
    for (int ky = kyl; ky <= kyh; ++ky) {
        const int64_t vy = (int64_t)sy + (int64_t)ky * ny;
        if (vy > ey - 1 || vy < -(int64_t)(ey - 1)) continue;
        const int64_t ay = vy < 0 ? -vy : vy;
        for (int kx = kxl; kx <= kxh; ++kx) {
            const int64_t vx = (int64_t)sx + (int64_t)kx * nx;
            if (vx > ex - 1 || vx < -(int64_t)(ex - 1)) continue;
            const int64_t ax = vx < 0 ? -vx : vx;
            const double *e = ext + (ay * ex + ax) * 4;
            const double sg = (vx == 0 || vy == 0) ? 0.0 : ((vx > 0) == (vy > 0) ? 1.0 : -1.0);
            acc[0] += e[0];
            acc[1] += sg * e[1];
            acc[2] += e[2];
            acc[3] += e[3];
        }
    }
    // base keeps the +x +y quadrant convention of k_quad (k_reflect signs kxy)
    base[t * 4 + 0] = acc[0];
    base[t * 4 + 1] = acc[1];
    base[t * 4 + 2] = acc[2];
    base[t * 4 + 3] = acc[3];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename T>
static cpx<T> *make_twiddles(int code) {
    // exp(-2 pi i t / L), t < L, for L = 2^code; for a 3M code the M-point
    // table followed by the 3M-point table
    std::vector<cpx<T>> h;
    auto add = [&](int L) {
        for (int t = 0; t < L; ++t) {
            const double a = -2.0 * M_PI * (double)t / (double)L;
            h.push_back({T(std::cos(a)), T(std::sin(a))});
        }
    };
    if (code_f3(code)) {
        add(1 << code_log2m(code));
        add(3 << code_log2m(code));
    } else {
        add(1 << code);
    }
    cpx<T> *d = nullptr;
    LLGMI_HIP(hipMalloc(&d, sizeof(cpx<T>) * h.size()));
    LLGMI_HIP(hipMemcpy(d, h.data(), sizeof(cpx<T>) * h.size(), hipMemcpyHostToDevice));
    return d;
}

static constexpr size_t kLdsMax = 160 * 1024;

// Dynamic LDS above the default per-launch limit must be opted into.
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
    if (bytes > kLdsMax) throw Error(3, "FFT block needs more LDS than a CU has");
    if (bytes > 64 * 1024)
        LLGMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Blocks of `kernel` (nt threads, lds bytes of dynamic LDS) resident on the
// current device at once: the grid of a persistent launch.  Cached per
// (kernel, shape, device).
static int resident_blocks(const void *kernel, int nt, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, size_t, int>, int> cache;
    int dev = 0;
    LLGMI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(kernel, nt, lds, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, ncu = 0;
    LLGMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, lds));
    LLGMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int nb = std::max(1, per_cu) * std::max(1, ncu);
    cache[key] = nb;
    return nb;
}

// A column block holds two sequences, 2 * L/R0 threads (at most 1024).
template <int LG, int RQ>
constexpr bool col_rq_ok() { return 2 * Plan<LG, RQ>::TPS <= 1024; }

// Built ranges (padded lengths, so nx, ny <= 4096 in FP32 and nx <= 4096,
// ny <= 2048 in FP64): rows up to 2^13 in both precisions; columns up to 2^13
// (FP32) / 2^12 (FP64).  A column block holds the two padded sequences of one
// item in LDS (2 * 8712 * 16 B = 272 KiB for an FP64 column of 2^13, beyond a
// CU's 160 KiB) and 2 * TPS <= 1024 threads.  The FP32 column of 2^13 spills
// a few hundred bytes per lane; it is correct but not tuned.
template <typename T> struct Range;
template <> struct Range<float> { static constexpr int ROW = 13, COL = 13; };
template <> struct Range<double> { static constexpr int ROW = 13, COL = 12; };
// Columns beyond Range<T>::COL, up to this length, take the split column pass
// (FP64 columns of 2^13; an FP32 column of 2^13 still fits the paired block).
static constexpr int kColSplitMax = 13;
template <typename T, int LG>
constexpr size_t colsplit_lds() { return sizeof(cpx<T>) * lds_seq_len(1 << LG); }
static constexpr int kRowF64Build = 13;  // FP64 row pass of the kernel build (FP32 contexts)
static constexpr int kExtractMax = 13;
// Lengths 3 * 2^k built per precision and pass (k = log2 M); the padding takes
// the shorter of the next power of two and the next 3 * 2^k (a grid of
// 3 * 2^(k-1) cells pads to 3 * 2^k, as the reference's 2n does).
template <typename T> struct Range3;
template <> struct Range3<float> { static constexpr int LO = 8, ROW = 11, COL = 11; };
template <> struct Range3<double> { static constexpr int LO = 8, ROW = 11, COL = 10; };
static constexpr int kF3BuildLo = 8, kF3BuildHi = 11;   // kernel build (FP64 rows and column extraction)
static constexpr int kF3AutoLo = 10;   // shortest 3 * 2^k taken without LLGMI_FFT_PAD=3
// dispatch over the length codes 1..HI (powers of two) and kF3 + LO3..kF3 + HI3
template <int HI, int LO3, int HI3, typename F>
inline void dispatch_code(int c, F &&f) {
    if (code_f3(c)) dispatch_log2<kF3 + LO3, kF3 + HI3>(c, static_cast<F &&>(f));
    else dispatch_log2<1, HI>(c, static_cast<F &&>(f));
}
template <typename T, typename F>
inline void dispatch_row(int c, F &&f) {
    dispatch_code<Range<T>::ROW, Range3<T>::LO, Range3<T>::ROW>(c, static_cast<F &&>(f));
}
template <typename T, typename F>
inline void dispatch_col(int c, F &&f) {
    dispatch_code<Range<T>::COL, Range3<T>::LO, Range3<T>::COL>(c, static_cast<F &&>(f));
}

template <typename T>
DemagFFT<T>::~DemagFFT() {
    release();
}

template <typename T>
void DemagFFT<T>::release() {
    void *ps[] = {tw_x, tw_y, kern, spec, colfull};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    tw_x = tw_y = nullptr;
    kern = nullptr;
    spec = nullptr;
    colfull = nullptr;
    built = false;
}

static int forced_radix_log2() {
    const char *f = std::getenv("LLGMI_FFT_RADIX");
    if (!f || !*f) return 0;
    const std::string v(f);
    if (v == "8") return 3;
    if (v == "16") return 4;
    throw Error(1, "LLGMI_FFT_RADIX must be 8 or 16, got '" + v + "'");
}

// Process-wide cache of real-space base quadrants keyed by grid and periodic
// images (the reference keeps its operators in a process-global cache keyed
// the same way, demag_fft_uniform.rs:44-108): a second context, or the other
// precision, on the same grid skips the quadrature.
namespace {
struct BaseKey {
    int64_t nx, ny, pbc_x, pbc_y;
    uint64_t dx, dy, dz;
    bool operator<(const BaseKey &o) const {
        return std::tie(nx, ny, pbc_x, pbc_y, dx, dy, dz) < std::tie(o.nx, o.ny, o.pbc_x, o.pbc_y, o.dx, o.dy, o.dz);
    }
};
std::mutex g_base_mu;
std::map<BaseKey, std::vector<double>> g_base_cache;
BaseKey base_key(int64_t nx, int64_t ny, double dx, double dy, double dz, int64_t px, int64_t py) {
    BaseKey k{nx, ny, px, py, 0, 0, 0};
    std::memcpy(&k.dx, &dx, 8);
    std::memcpy(&k.dy, &dy, 8);
    std::memcpy(&k.dz, &dz, 8);
    return k;
}

// Tuned kernel choices, shared process-wide by every operator of one padded
// shape and precision on one device: the first operator times the variants,
// later ones (another context, the films of a batch) take its choice, so that
// operators of the same shape always run the same kernels (their results are
// then bitwise alike) and the timing is paid once.
struct TuneKey {
    int prec, dev;
    int64_t nx, ny, px, py;
    bool operator<(const TuneKey &o) const {
        return std::tie(prec, dev, nx, ny, px, py) < std::tie(o.prec, o.dev, o.nx, o.ny, o.px, o.py);
    }
};
struct TuneVal {
    int rq_row = 0, rq_col = 0, rq_row_inv = 0;   // 0: not tuned yet
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
// the fused-inverse choice, by shape and the unfused inverse pass's radix
std::map<std::pair<TuneKey, int>, int> g_tune_fused;
template <typename T>
TuneKey tune_key(const DemagFFT<T> &op) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    return TuneKey{(int)sizeof(T), dev, op.nx, op.ny, op.px, op.py};
}
}  // namespace

template <typename T>
void DemagFFT<T>::build(int64_t nx_, int64_t ny_, double dx_, double dy_, double dz_,
                        hipStream_t stream, int64_t pbc_x_, int64_t pbc_y_,
                        const std::vector<double> *base_in) {
    std::vector<double> cached;
    const BaseKey bkey = base_key(nx_, ny_, dx_, dy_, dz_, pbc_x_, pbc_y_);
    if (!base_in) {
        std::lock_guard<std::mutex> lock(g_base_mu);
        auto it = g_base_cache.find(bkey);
        if (it != g_base_cache.end()) {
            cached = it->second;
            base_in = &cached;
        }
    }
    const int forced = forced_radix_log2();
    release();
    nx = nx_; ny = ny_; dx = dx_; dy = dy_; dz = dz_;
    pbc_x = pbc_x_; pbc_y = pbc_y_;
    // kernel range per axis: |s| <= n pbc - 1 on a periodic axis, n - 1 on an open one
    const int64_t ex = pbc_x > 0 ? nx * pbc_x : nx, ey = pbc_y > 0 ? ny * pbc_y : ny;
    if (pbc_x < 0 || pbc_y < 0 || pbc_x > (1 << 20) || pbc_y > (1 << 20) || ex * ey > (int64_t(1) << 27))
        throw Error(3, "FFT demag: periodic image range " + std::to_string(ex) + " x " + std::to_string(ey) +
                           " cells is beyond the built limit of 2^27 kernel displacements");
    // padded lengths >= 2n: the next power of two, or 3 * 2^k when that is
    // shorter, built and at least 3072 long (at 1536 the three-transform form
    // measured slower than the power of two it replaces: 61.7 against 53.7 µs
    // per 768^2 FP32 evaluation; at 3072 faster: 141 against 170 µs at 1536^2).
    // LLGMI_FFT_PAD=pow2 keeps powers of two, =3 takes every shorter 3 * 2^k.
    const char *pad = std::getenv("LLGMI_FFT_PAD");
    const std::string pads = pad ? pad : "";
    if (!pads.empty() && pads != "pow2" && pads != "3")
        throw Error(1, "LLGMI_FFT_PAD must be pow2 or 3, got '" + pads + "'");
    const int lo3 = pads == "3" ? Range3<T>::LO : kF3AutoLo;
    auto pick = [&](int64_t need, int hi3) {
        const int64_t p2 = next_pow2(need);
        int code = ilog2(p2);
        if (pads != "pow2")
            for (int k = lo3; k <= hi3; ++k)
                if ((int64_t(3) << k) >= need) {
                    if ((int64_t(3) << k) < p2) code = kF3 + k;
                    break;
                }
        return code;
    };
    lx = pick(2 * nx, Range3<T>::ROW);
    ly = pick(2 * ny, Range3<T>::COL);
    px = code_len(lx);
    py = code_len(ly);
    hx = px / 2 + 1;
    hy = py / 2 + 1;
    if ((!code_f3(lx) && lx > Range<T>::ROW) || (!code_f3(ly) && (ly > kColSplitMax || ly > kExtractMax)))
        throw Error(3, "FFT demag: padded size " + std::to_string(next_pow2(2 * nx)) + "x" +
                           std::to_string(next_pow2(2 * ny)) + " is beyond the built FFT lengths of this precision");
    const double mu0 = 4.0 * M_PI * 1e-7;
    const double acc = 10.0;  // DEMAG_ACCURACY (demag_fft_uniform.rs:432)

    double *base = nullptr, *kr = nullptr, *ext = nullptr;
    cpx<double> *sk = nullptr, *twx64 = nullptr, *twy64 = nullptr;
    try {
        // 1) base-quadrant quadrature (FP64); with periodic axes over the whole
        //    image range, then folded onto the bins
        const int64_t nb = nx * ny;
        LLGMI_HIP(hipMalloc(&base, sizeof(double) * 4 * nb));
        const bool periodic = (pbc_x > 0 || pbc_y > 0) && !base_in;
        if (base_in) {
            if ((int64_t)base_in->size() != 4 * nb) throw Error(1, "kernel cache: base quadrant of the wrong size");
            LLGMI_HIP(hipMemcpyAsync(base, base_in->data(), sizeof(double) * 4 * nb, hipMemcpyHostToDevice, stream));
        }
        if (periodic) LLGMI_HIP(hipMalloc(&ext, sizeof(double) * 4 * ex * ey));
        if (!base_in) {
            int nt = 256;
            int64_t nq = periodic ? ex * ey : nb;
            int64_t nbk = (3 * nq + nt - 1) / nt;
            const int qnx = (int)(periodic ? ex : nx), qny = (int)(periodic ? ey : ny);
            hipLaunchKernelGGL(k_quad, dim3((unsigned)nbk), dim3(nt), 0, stream, periodic ? ext : base, qnx, qny,
                               dx, dy, dz, acc, mu0);
            LLGMI_LAUNCH_CHECK();
            // heavy near-field sums: the point counts fall with distance, so they
            // sit in a small corner of the quadrant; found on the host
            std::vector<int> items;
            const int lim = kQuadHeavyScan;
            for (int sy = 0; sy < std::min(qny, lim); ++sy)
                for (int sx = 0; sx < std::min(qnx, lim); ++sx)
                    for (int u = 0; u < 3; ++u) {
                        const QuadSetup q = quad_setup(u, dx, dy, dz, sx, sy, acc);
                        if (q.nv * q.nw * q.nx * q.ny * q.nz > kQuadHeavy) items.push_back((sy * qnx + sx) * 3 + u);
                    }
            if (!items.empty()) {
                int *d_items = nullptr;
                LLGMI_HIP(hipMalloc(&d_items, sizeof(int) * items.size()));
                hipError_t e1 = hipMemcpyAsync(d_items, items.data(), sizeof(int) * items.size(),
                                               hipMemcpyHostToDevice, stream);
                if (e1 == hipSuccess) {
                    hipLaunchKernelGGL(k_quad_heavy, dim3((unsigned)items.size()), dim3(kQuadChunk), 0, stream,
                                       periodic ? ext : base, d_items, qnx, dx, dy, dz, acc, mu0);
                    e1 = hipGetLastError();
                }
                hipError_t e2 = hipStreamSynchronize(stream);
                (void)hipFree(d_items);
                LLGMI_HIP(e1);
                LLGMI_HIP(e2);
            }
        }
        if (periodic) {
            hipLaunchKernelGGL(k_fold_images, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream, ext,
                               (int)ex, (int)ey, (int)nx, (int)ny, (int)pbc_x, (int)pbc_y, base);
            LLGMI_LAUNCH_CHECK();
        }
        base_host.resize((size_t)(4 * nb));
        LLGMI_HIP(hipMemcpyAsync(base_host.data(), base, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, stream));
        // 2) reflect into padded real arrays
        const int64_t np = px * py;
        LLGMI_HIP(hipMalloc(&kr, sizeof(double) * 4 * np));
        hipLaunchKernelGGL(k_reflect, dim3((unsigned)((px + 255) / 256), (unsigned)py), dim3(256), 0,
                           stream, base, (int)nx, (int)ny, (int)px, (int)py, kr);
        LLGMI_LAUNCH_CHECK();
        // 3) row R2C (FP64) of the four arrays
        twx64 = make_twiddles<double>(lx);
        twy64 = make_twiddles<double>(ly);
        const int64_t hxs = px / 2;
        LLGMI_HIP(hipMalloc(&sk, sizeof(cpx<double>) * 4 * hxs * py));
        dispatch_code<kRowF64Build, kF3BuildLo, kF3BuildHi>(lx, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            using C = RowCfg<double, LG, 0>;
            allow_lds(k_row_r2c<double, double, LG, false, 0>, C::LDS);
            dim3 grid((unsigned)((py + 2 * C::PAIRS - 1) / (2 * C::PAIRS)), 4);
            hipLaunchKernelGGL((k_row_r2c<double, double, LG, false, 0>), grid, dim3(C::NT), C::LDS, stream, kr,
                               np, (int)px, (int)px, (int)py, 1.0, twx64, sk, (int64_t)hxs * py);
            LLGMI_LAUNCH_CHECK();
        });
        // 4) column FFT (FP64), real quadrant -> kern (T)
        LLGMI_HIP(hipMalloc(&kern, sizeof(T) * 4 * hx * hy));
        dispatch_code<kExtractMax, kF3BuildLo, kF3BuildHi>(ly, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            using C = ExtractCfg<LG>;
            allow_lds(k_col_extract<T, LG>, C::LDS);
            hipLaunchKernelGGL((k_col_extract<T, LG>), dim3((unsigned)hxs, 4), dim3(C::NT), C::LDS,
                               stream, sk, (int)hxs, twy64, kern, 1.0 / ((double)px * (double)py));
            LLGMI_LAUNCH_CHECK();
        });
        LLGMI_HIP(hipStreamSynchronize(stream));
    } catch (...) {
        if (base) (void)hipFree(base);
        if (ext) (void)hipFree(ext);
        if (kr) (void)hipFree(kr);
        if (sk) (void)hipFree(sk);
        if (twx64) (void)hipFree(twx64);
        if (twy64) (void)hipFree(twy64);
        release();
        throw;
    }
    (void)hipFree(base);
    if (ext) (void)hipFree(ext);
    (void)hipFree(kr);
    (void)hipFree(sk);
    {
        std::lock_guard<std::mutex> lock(g_base_mu);
        if (g_base_cache.size() >= 16) g_base_cache.clear();
        g_base_cache[bkey] = base_host;
    }
    (void)hipFree(twx64);
    (void)hipFree(twy64);

    tw_x = make_twiddles<T>(lx);
    tw_y = make_twiddles<T>(ly);
    LLGMI_HIP(hipMalloc(&spec, sizeof(cpx<T>) * 3 * (px / 2) * ny));
    if (col_split()) LLGMI_HIP(hipMalloc(&colfull, sizeof(cpx<T>) * 3 * (px / 2) * py));
    dispatch_row<T>(lx, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        allow_lds(k_row_r2c<T, T, LG, true, 3>, RowCfg<T, LG, 3>::LDS);
        allow_lds(k_row_r2c<T, T, LG, false, 3>, RowCfg<T, LG, 3>::LDS);
        allow_lds(k_row_c2r<T, LG, true, 3>, RowCfg<T, LG, 3>::LDS);
        allow_lds(k_row_c2r<T, LG, false, 3>, RowCfg<T, LG, 3>::LDS);
        allow_lds(k_row_r2c<T, T, LG, true, 4>, RowCfg<T, LG, 4>::LDS);
        allow_lds(k_row_r2c<T, T, LG, false, 4>, RowCfg<T, LG, 4>::LDS);
        allow_lds(k_row_c2r<T, LG, true, 4>, RowCfg<T, LG, 4>::LDS);
        allow_lds(k_row_c2r<T, LG, false, 4>, RowCfg<T, LG, 4>::LDS);
    });
    dispatch_row<T>(lx, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        auto allow = [&](auto rqc) {
            constexpr int RQ = decltype(rqc)::value;
            if constexpr (fused_ok<T, LG, RQ>()) {
                constexpr size_t lds = FusedCfg<T, LG, RQ>::LDS;
                allow_lds(k_row_c2r_local<T, LG, true, RQ, false, false>, lds);
                allow_lds(k_row_c2r_local<T, LG, true, RQ, false, true>, lds);
                allow_lds(k_row_c2r_local<T, LG, true, RQ, true, false>, lds);
                allow_lds(k_row_c2r_local<T, LG, true, RQ, true, true>, lds);
                allow_lds(k_row_c2r_local<T, LG, false, RQ, false, false>, lds);
                allow_lds(k_row_c2r_local<T, LG, false, RQ, false, true>, lds);
                allow_lds(k_row_c2r_local<T, LG, false, RQ, true, false>, lds);
                allow_lds(k_row_c2r_local<T, LG, false, RQ, true, true>, lds);
            }
        };
        allow(std::integral_constant<int, 3>{});
        allow(std::integral_constant<int, 4>{});
    });
    if (col_split()) {
        dispatch_log2<Range<T>::COL + 1, kColSplitMax>(ly, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            allow_lds(k_colsplit_fwd<T, LG, true, 3>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_fwd<T, LG, false, 3>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_fwd<T, LG, true, 4>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_fwd<T, LG, false, 4>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_mul_inv<T, LG, true, 3>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_mul_inv<T, LG, false, 3>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_mul_inv<T, LG, true, 4>, colsplit_lds<T, LG>());
            allow_lds(k_colsplit_mul_inv<T, LG, false, 4>, colsplit_lds<T, LG>());
        });
    } else dispatch_col<T>(ly, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        constexpr bool pipe_ok = !code_f3(LG);   // the persistent form is built for powers of two
        if constexpr (col_rq_ok<LG, 3>()) {
            allow_lds(k_col_demag<T, LG, true, 3, false>, ColGrp<T, LG, 3>::LDS);
            allow_lds(k_col_demag<T, LG, false, 3, false>, ColGrp<T, LG, 3>::LDS);
            if constexpr (pipe_ok) {
                allow_lds(k_col_demag<T, LG, true, 3, true>, ColCfg<T, LG, 3>::LDS);
                allow_lds(k_col_demag<T, LG, false, 3, true>, ColCfg<T, LG, 3>::LDS);
            }
        }
        allow_lds(k_col_demag<T, LG, true, 4, false>, ColGrp<T, LG, 4>::LDS);
        allow_lds(k_col_demag<T, LG, false, 4, false>, ColGrp<T, LG, 4>::LDS);
        if constexpr (pipe_ok) {
            allow_lds(k_col_demag<T, LG, true, 4, true>, ColCfg<T, LG, 4>::LDS);
            allow_lds(k_col_demag<T, LG, false, 4, true>, ColCfg<T, LG, 4>::LDS);
        }
    });
    built = true;
    autotune(stream, forced);
}

template <typename T>
bool DemagFFT<T>::col_split() const {
    return !code_f3(ly) && ly > Range<T>::COL;
}

template <typename T>
void DemagFFT<T>::launch_row_fwd(const T *m, T ms, hipStream_t stream, int rq) {
    const int64_t n = nx * ny;
    dispatch_row<T>(lx, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        auto go = [&](auto rqc) {
            constexpr int RQ = decltype(rqc)::value;
            using C = RowCfg<T, LG, RQ>;
            const dim3 grid((unsigned)((ny + 2 * C::PAIRS - 1) / (2 * C::PAIRS)), 3);
            if (2 * nx == px)
                hipLaunchKernelGGL((k_row_r2c<T, T, LG, true, RQ>), grid, dim3(C::NT), C::LDS, stream,
                                   m, n, (int)nx, (int)nx, (int)ny, ms, tw_x, spec, (int64_t)(px / 2) * ny);
            else
                hipLaunchKernelGGL((k_row_r2c<T, T, LG, false, RQ>), grid, dim3(C::NT), C::LDS, stream,
                                   m, n, (int)nx, (int)nx, (int)ny, ms, tw_x, spec, (int64_t)(px / 2) * ny);
        };
        if (rq == 3) go(std::integral_constant<int, 3>{});
        else go(std::integral_constant<int, 4>{});
        LLGMI_LAUNCH_CHECK();
    });
}

template <typename T>
void DemagFFT<T>::launch_col(hipStream_t stream, int rq) {
    if (col_split()) {
        // split column pass (the persistent flag does not apply)
        dispatch_log2<Range<T>::COL + 1, kColSplitMax>(ly, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            const int hxs = (int)(px / 2);
            const bool half = 2 * ny == py;
            const dim3 grid((unsigned)(3 * hxs));
            constexpr size_t lds = colsplit_lds<T, LG>();
            auto go = [&](auto rqc) {
                constexpr int RQ = decltype(rqc)::value;
                constexpr int NT = RegPlan<LG, RQ>::TPS;
                if (half) {
                    hipLaunchKernelGGL((k_colsplit_fwd<T, LG, true, RQ>), grid, dim3(NT), lds, stream, spec, colfull,
                                       (int)ny, tw_y);
                    hipLaunchKernelGGL((k_colsplit_mul_inv<T, LG, true, RQ>), grid, dim3(NT), lds, stream, spec,
                                       colfull, hxs, (int)ny, tw_y, kern);
                } else {
                    hipLaunchKernelGGL((k_colsplit_fwd<T, LG, false, RQ>), grid, dim3(NT), lds, stream, spec, colfull,
                                       (int)ny, tw_y);
                    hipLaunchKernelGGL((k_colsplit_mul_inv<T, LG, false, RQ>), grid, dim3(NT), lds, stream, spec,
                                       colfull, hxs, (int)ny, tw_y, kern);
                }
            };
            if ((rq & 7) == 3) go(std::integral_constant<int, 3>{});
            else go(std::integral_constant<int, 4>{});
            LLGMI_LAUNCH_CHECK();
        });
        return;
    }
    dispatch_col<T>(ly, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        // rq: log2 of the first/last radix, plus kColPipe for persistent blocks
        const bool pipe = (rq & kColPipe) != 0;
        auto go = [&](auto rqc) {
            constexpr int RQ = decltype(rqc)::value;
            using C = ColCfg<T, LG, RQ>;
            const int hxs = (int)(px / 2);
            const int nwork = hxs + (hxs + 1) / 2;  // (x, y) items + z-pair items
            const bool half = 2 * ny == py;
            bool launched = false;
            if constexpr (!code_f3(LG)) {   // the persistent form is built for powers of two
                if (pipe) {
                    auto kh = half ? k_col_demag<T, LG, true, RQ, true> : k_col_demag<T, LG, false, RQ, true>;
                    const int nb = std::min(nwork, resident_blocks((const void *)kh, C::NT, C::LDS));
                    hipLaunchKernelGGL(kh, dim3((unsigned)nb), dim3(C::NT), C::LDS, stream, spec, hxs, (int)ny,
                                       tw_y, kern, nwork);
                    launched = true;
                }
            }
            if (!launched) {
                using Gc = ColGrp<T, LG, RQ>;
                auto kh = half ? k_col_demag<T, LG, true, RQ, false> : k_col_demag<T, LG, false, RQ, false>;
                hipLaunchKernelGGL(kh, dim3((unsigned)((nwork + Gc::G - 1) / Gc::G)), dim3(Gc::NT), Gc::LDS, stream,
                                   spec, hxs, (int)ny, tw_y, kern, nwork);
            }
        };
        if constexpr (col_rq_ok<LG, 3>()) {
            if ((rq & 7) == 3) go(std::integral_constant<int, 3>{});
            else go(std::integral_constant<int, 4>{});
        } else {
            go(std::integral_constant<int, 4>{});
        }
        LLGMI_LAUNCH_CHECK();
    });
}

template <typename T>
void DemagFFT<T>::launch_row_inv(T *bd, hipStream_t stream, int rq) {
    const int64_t n = nx * ny;
    dispatch_row<T>(lx, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        auto go = [&](auto rqc) {
            constexpr int RQ = decltype(rqc)::value;
            using C = RowCfg<T, LG, RQ>;
            const dim3 grid((unsigned)((ny + 2 * C::PAIRS - 1) / (2 * C::PAIRS)), 3);
            auto launch = [&](auto kfn) {
                hipLaunchKernelGGL(kfn, grid, dim3(C::NT), C::LDS, stream, spec, (int)(px / 2), (int)nx, (int)ny,
                                   tw_x, bd, n);
            };
            launch(2 * nx == px ? k_row_c2r<T, LG, true, RQ> : k_row_c2r<T, LG, false, RQ>);
        };
        if (rq == 3) go(std::integral_constant<int, 3>{});
        else go(std::integral_constant<int, 4>{});
        LLGMI_LAUNCH_CHECK();
    });
}

// Per grid, time the radix-8 and radix-16 variants of each pass and keep the
// faster: which one wins depends on the lengths through occupancy (registers)
// versus LDS round trips.  LLGMI_FFT_RADIX=8 or 16 forces one variant for all
// passes (used by tests to cover both); grids below 256 x 256 padded points
// are launch-bound and keep radix 16 without timing.
template <typename T>
void DemagFFT<T>::autotune(hipStream_t stream, int forced) {
    // LLGMI_FFT_COLPIPE=0|1 forces the one-item-per-block or the persistent
    // column pass
    const char *pe = std::getenv("LLGMI_FFT_COLPIPE");
    const int pipe_forced = pe ? (pe[0] == '1' ? 1 : 0) : -1;
    rq_row = rq_col = rq_row_inv = forced ? forced : 4;
    if (pipe_forced == 1) rq_col |= kColPipe;
    if (forced || px * py < 256 * 256) return;
    const bool shared = pipe_forced < 0;   // forced settings neither read nor set the shared choice
    const TuneKey key = tune_key(*this);
    if (shared) {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end() && it->second.rq_row) {
            rq_row = it->second.rq_row;
            rq_col = it->second.rq_col;
            rq_row_inv = it->second.rq_row_inv;
            return;
        }
    }
    const int64_t n = nx * ny;
    T *m = nullptr, *bd = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (m) (void)hipFree(m);
        if (bd) (void)hipFree(bd);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        LLGMI_HIP(hipMalloc(&m, sizeof(T) * 3 * n));
        LLGMI_HIP(hipMalloc(&bd, sizeof(T) * 3 * n));
        LLGMI_HIP(hipMemsetAsync(m, 0, sizeof(T) * 3 * n, stream));
        LLGMI_HIP(hipEventCreate(&e0));
        LLGMI_HIP(hipEventCreate(&e1));
        auto time = [&](auto &&launch) {
            LLGMI_HIP(hipEventRecord(e0, stream));
            for (int i = 0; i < 4; ++i) launch();
            LLGMI_HIP(hipEventRecord(e1, stream));
            LLGMI_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LLGMI_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        // The three passes are timed together for all radix / column-pass
        // combinations (a pass alone sees other cache contents than inside the
        // pipeline); combinations alternate over five rounds, in reverse order
        // every other round, and the minimum per combination decides.  A
        // warm-up of about 20 ms first brings the clocks up, so that no
        // combination profits from coming later.
        auto pipeline = [&](int rr, int rc, int ri) {
            launch_row_fwd(m, T(1), stream, rr);
            launch_col(stream, rc);
            launch_row_inv(bd, stream, ri);
        };
        {
            const float t1 = time([&] { pipeline(4, 4, 4); }) / 4;
            const int reps = std::min(2000, std::max(8, (int)(20.f / std::max(t1, 1e-3f))));
            for (int i = 0; i < reps; ++i) pipeline(4, 4, 4);
            LLGMI_HIP(hipStreamSynchronize(stream));
        }
        float best[16];
        for (float &b : best) b = 1e30f;
        auto allowed = [&](int c) {
            const int p = (c >> 3) & 1;
            return pipe_forced < 0 || p == pipe_forced;
        };
        for (int round = 0; round < 5; ++round)
            for (int i = 0; i < 16; ++i) {
                const int c = (round & 1) ? 15 - i : i;
                if (!allowed(c)) continue;
                const int rr = 3 + (c & 1), rc = (3 + ((c >> 1) & 1)) | ((c & 8) ? kColPipe : 0),
                          ri = 3 + ((c >> 2) & 1);
                const float t = time([&] { pipeline(rr, rc, ri); });
                if (t < best[c]) best[c] = t;
            }
        int bc = -1;
        for (int c = 0; c < 16; ++c)
            if (allowed(c) && (bc < 0 || best[c] < best[bc])) bc = c;
        rq_row = 3 + (bc & 1);
        rq_col = (3 + ((bc >> 1) & 1)) | ((bc & 8) ? kColPipe : 0);
        rq_row_inv = 3 + ((bc >> 2) & 1);
        if (shared) {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            TuneVal &v = g_tune[key];
            v.rq_row = rq_row;
            v.rq_col = rq_col;
            v.rq_row_inv = rq_row_inv;
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

template <typename T>
void DemagFFT<T>::field(const T *m, T ms, T *bd, hipStream_t stream, Profiler *prof) {
    {
        ProfScope ps(prof, 0, stream);
        launch_row_fwd(m, ms, stream, rq_row);
    }
    {
        ProfScope ps(prof, 1, stream);
        launch_col(stream, rq_col);
    }
    {
        ProfScope ps(prof, 2, stream);
        launch_row_inv(bd, stream, rq_row_inv);
    }
}

template <typename T>
void DemagFFT<T>::launch_row_inv_local(const LocalArgs<T> &a, const T *m, const uint8_t *geom, T *b, T *kout,
                                       hipStream_t stream, int rq) {
    dispatch_row<T>(lx, [&](auto lg) {
        constexpr int LG = decltype(lg)::value;
        auto go = [&](auto rqc) {
            constexpr int RQ = decltype(rqc)::value;
            if constexpr (fused_ok<T, LG, RQ>()) {
                using C = FusedCfg<T, LG, RQ>;
                const dim3 grid((unsigned)((ny + 2 * C::PAIRS - 1) / (2 * C::PAIRS)));
                const bool half = 2 * nx == px, g = geom != nullptr, q = kout != nullptr;
#define LLGMI_FUSED(H, G, Q)                                                                          \
    hipLaunchKernelGGL((k_row_c2r_local<T, LG, H, RQ, G, Q>), grid, dim3(C::NT), C::LDS, stream, spec, \
                       (int)(px / 2), (int)ny, tw_x, a, m, geom, b, kout)
                if (half) {
                    if (g) { if (q) LLGMI_FUSED(true, true, true); else LLGMI_FUSED(true, true, false); }
                    else { if (q) LLGMI_FUSED(true, false, true); else LLGMI_FUSED(true, false, false); }
                } else {
                    if (g) { if (q) LLGMI_FUSED(false, true, true); else LLGMI_FUSED(false, true, false); }
                    else { if (q) LLGMI_FUSED(false, false, true); else LLGMI_FUSED(false, false, false); }
                }
#undef LLGMI_FUSED
                LLGMI_LAUNCH_CHECK();
            } else {
                throw Error(3, "fused inverse row pass: block too large for this radix");
            }
        };
        if (rq == 3) go(std::integral_constant<int, 3>{});
        else go(std::integral_constant<int, 4>{});
    });
}

// First call on a grid: the fused pass (radix 8 or 16, where the block fits)
// and the unfused pair (inverse row pass + local-field kernel) are timed on
// the real arguments and the fastest kept.  LLGMI_FFT_FUSED=0 / 1 forces the
// unfused / fused path (fused with the radix of the inverse row pass).
template <typename T>
void DemagFFT<T>::field_local(const T *m, T ms, const LocalArgs<T> &a, const uint8_t *geom, T *b, T *kout,
                              T *bd_scratch, hipStream_t stream, Profiler *prof) {
    {
        ProfScope ps(prof, 0, stream);
        launch_row_fwd(m, ms, stream, rq_row);
    }
    {
        ProfScope ps(prof, 1, stream);
        launch_col(stream, rq_col);
    }
    auto fits = [&](int rq) {
        bool ok = false;
        dispatch_row<T>(lx, [&](auto lg) {
            constexpr int LG = decltype(lg)::value;
            ok = rq == 3 ? fused_ok<T, LG, 3>() : fused_ok<T, LG, 4>();
        });
        return ok;
    };
    auto unfused = [&] {
        launch_row_inv(bd_scratch, stream, rq_row_inv);
        launch_local_field<T>(a, m, bd_scratch, geom, b, kout, stream);
    };
    if (rq_fused == 0) {
        const char *f = std::getenv("LLGMI_FFT_FUSED");
        const std::string fv = f ? std::string(f) : std::string();
        if (fv == "0") {
            rq_fused = -1;
        } else if (fv == "1") {
            rq_fused = fits(rq_row_inv) ? rq_row_inv : (fits(3) ? 3 : (fits(4) ? 4 : -1));
        } else if (px * py < 256 * 256) {
            rq_fused = fits(4) ? 4 : (fits(3) ? 3 : -1);
        } else {
            const std::pair<TuneKey, int> fkey{tune_key(*this), rq_row_inv};
            {
                std::lock_guard<std::mutex> lock(g_tune_mu);
                auto it = g_tune_fused.find(fkey);
                if (it != g_tune_fused.end()) rq_fused = it->second;
            }
        }
        if (rq_fused == 0) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            try {
                LLGMI_HIP(hipEventCreate(&e0));
                LLGMI_HIP(hipEventCreate(&e1));
                auto time = [&](auto &&launch) {
                    LLGMI_HIP(hipEventRecord(e0, stream));
                    for (int i = 0; i < 4; ++i) launch();
                    LLGMI_HIP(hipEventRecord(e1, stream));
                    LLGMI_HIP(hipEventSynchronize(e1));
                    float ms_ = 0.f;
                    LLGMI_HIP(hipEventElapsedTime(&ms_, e0, e1));
                    return ms_;
                };
                float best[3] = {1e30f, 1e30f, 1e30f};   // unfused, fused radix 8, fused radix 16
                auto run = [&](int v) {
                    if (v == 0) best[0] = std::min(best[0], time(unfused));
                    else if (fits(v == 1 ? 3 : 4))
                        best[v] = std::min(best[v], time([&] { launch_row_inv_local(a, m, geom, b, kout, stream, v == 1 ? 3 : 4); }));
                };
                for (int round = 0; round < 5; ++round)   // in reverse order every other round
                    for (int i = 0; i < 3; ++i) run((round & 1) ? 2 - i : i);
                // the fused kernel must win by 3 %: within that the two are a
                // tie on the plain field, and the unfused pair is the faster
                // one once the torque and the next RK stage ride on the local
                // field (the fused kernel's one block per CU leaves no slack
                // for them)
                const float fb = std::min(best[1], best[2]);
                rq_fused = fb < 0.97f * best[0] ? (best[1] <= best[2] ? 3 : 4) : -1;
                std::lock_guard<std::mutex> lock(g_tune_mu);
                g_tune_fused[{tune_key(*this), rq_row_inv}] = rq_fused;
            } catch (...) {
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                throw;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
    ProfScope ps(prof, 2, stream);
    if (rq_fused < 0) unfused();
    else launch_row_inv_local(a, m, geom, b, kout, stream, rq_fused);
}

// ---------------------------------------------------------------------------
// Batched passes: the block of table entry blockIdx.x runs the single-film
// body for its film and block.  One launch covers the films of every length
// in [LO, HI] whose blocks have the same shape (threads per block); a block
// runs the body of its own film's length (the kernel takes the largest LDS
// of the range).
// ---------------------------------------------------------------------------
template <typename T, int KIND, int LG, int RQ>
struct BatchCfg;
template <typename T, int LG, int RQ>
struct BatchCfg<T, 0, LG, RQ> {
    static constexpr bool OK = LG >= 1 && LG <= Range<T>::ROW;
    static constexpr int NT = RowCfg<T, LG, RQ>::NT;
    static constexpr size_t LDS = RowCfg<T, LG, RQ>::LDS;
};
template <typename T, int LG, int RQ>
struct BatchCfg<T, 2, LG, RQ> : BatchCfg<T, 0, LG, RQ> {};
template <typename T, int LG, int RQ>
struct BatchCfg<T, 1, LG, RQ> {
    static constexpr bool OK = LG >= 1 && LG <= Range<T>::COL && (RQ == 4 || col_rq_ok<LG, 3>());
    static constexpr int NT = ColGrp<T, LG, RQ>::NT;
    static constexpr size_t LDS = ColGrp<T, LG, RQ>::LDS;
};
template <typename T, int LG, int RQ>
struct BatchCfg<T, 3, LG, RQ> {
    static constexpr bool OK = LG >= 1 && LG <= Range<T>::ROW && BatchFusedCfg<T, LG, RQ>::OK;
    static constexpr int NT = BatchFusedCfg<T, LG, RQ>::NT;
    static constexpr size_t LDS = BatchFusedCfg<T, LG, RQ>::LDS;
};
template <typename T, int KIND>
constexpr int batch_max_lg() { return KIND == 1 ? Range<T>::COL : Range<T>::ROW; }

// the range of lengths around LG whose blocks have LG's shape
template <typename T, int KIND, int RQ, int LG>
constexpr int batch_run_lo() {
    if constexpr (LG > 1) {
        if constexpr (BatchCfg<T, KIND, LG - 1, RQ>::OK &&
                      BatchCfg<T, KIND, LG - 1, RQ>::NT == BatchCfg<T, KIND, LG, RQ>::NT)
            return batch_run_lo<T, KIND, RQ, LG - 1>();
        else
            return LG;
    } else {
        return LG;
    }
}
template <typename T, int KIND, int RQ, int LG>
constexpr int batch_run_hi() {
    if constexpr (LG < batch_max_lg<T, KIND>()) {
        if constexpr (BatchCfg<T, KIND, LG + 1, RQ>::OK &&
                      BatchCfg<T, KIND, LG + 1, RQ>::NT == BatchCfg<T, KIND, LG, RQ>::NT)
            return batch_run_hi<T, KIND, RQ, LG + 1>();
        else
            return LG;
    } else {
        return LG;
    }
}
template <typename T, int KIND, int RQ, int LO, int HI>
constexpr size_t batch_run_lds() {
    if constexpr (LO <= HI) {
        constexpr size_t a = BatchCfg<T, KIND, LO, RQ>::LDS, b = batch_run_lds<T, KIND, RQ, LO + 1, HI>();
        return a > b ? a : b;
    } else {
        return 0;
    }
}

// the body for length lg of [LO, HI] (lg is uniform over the block)
template <int LO, int HI, typename Fn>
__device__ __forceinline__ void dev_dispatch_log2(int lg, Fn &&fn) {
    if constexpr (LO <= HI) {
        if (lg == LO) fn(std::integral_constant<int, LO>{});
        else dev_dispatch_log2<LO + 1, HI>(lg, static_cast<Fn &&>(fn));
    }
}

template <typename T, int LO, int HI, bool HALF, int RQ>
__global__ void __launch_bounds__((RowCfg<T, LO, RQ>::NT))
kb_row_r2c(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int which, int stage1) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (stage1 && f.skip1) return;
    dev_dispatch_log2<LO, HI>(__builtin_ctz((unsigned)f.px), [&](auto lgc) {
        constexpr int LG = decltype(lgc)::value;
        constexpr int PAIRS = RowCfg<T, LG, RQ>::PAIRS;
        const int nb = (f.ny + 2 * PAIRS - 1) / (2 * PAIRS);
        const int comp = e.blk / nb, j0 = (e.blk - comp * nb) * 2 * PAIRS;
        row_r2c_body<T, T, LG, HALF, RQ>(f.in[which], (int64_t)f.nx * f.ny, f.nx, f.nx, f.ny, f.ms, f.tw_x, f.spec,
                                         (int64_t)(f.px / 2) * f.ny, comp, j0);
    });
}

template <typename T, int LO, int HI, bool HALF, int RQ>
__global__ void __launch_bounds__((ColGrp<T, LO, RQ>::NT))
kb_col(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int stage1) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (stage1 && f.skip1) return;
    const int hxs = f.px / 2;
    dev_dispatch_log2<LO, HI>(__builtin_ctz((unsigned)f.py), [&](auto lgc) {
        constexpr int LG = decltype(lgc)::value;
        col_work<T, LG, HALF, RQ>(f.spec, hxs, f.ny, f.tw_y, f.kern, hxs + (hxs + 1) / 2, e.blk);
    });
}

template <typename T, int LO, int HI, bool HALF, int RQ>
__global__ void __launch_bounds__((RowCfg<T, LO, RQ>::NT))
kb_row_c2r(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int stage1) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (stage1 && f.skip1) return;
    dev_dispatch_log2<LO, HI>(__builtin_ctz((unsigned)f.px), [&](auto lgc) {
        constexpr int LG = decltype(lgc)::value;
        constexpr int PAIRS = RowCfg<T, LG, RQ>::PAIRS;
        const int nb = (f.ny + 2 * PAIRS - 1) / (2 * PAIRS);
        const int comp = e.blk / nb, j0 = (e.blk - comp * nb) * 2 * PAIRS;
        row_c2r_body<T, LG, HALF, RQ>(f.spec, f.px / 2, f.nx, f.ny, f.tw_x, f.bd, (int64_t)f.nx * f.ny, comp, j0);
    });
}

template <typename T, int LO, int HI, bool HALF, int RQ>
__global__ void __launch_bounds__((BatchFusedCfg<T, LO, RQ>::NT))
kb_row_c2r_local(const BFilm<T> *__restrict__ F, const BEntry *__restrict__ tab, int which, int ks, int stage1) {
    const BEntry e = tab[blockIdx.x];
    const BFilm<T> &f = F[e.film];
    if (stage1 && f.skip1) return;
    dev_dispatch_log2<LO, HI>(__builtin_ctz((unsigned)f.px), [&](auto lgc) {
        constexpr int LG = decltype(lgc)::value;
        using C = BatchFusedCfg<T, LG, RQ>;
        row_c2r_local_body<T, LG, HALF, RQ, false, true, C>(f.spec, f.px / 2, f.ny, f.tw_x, f.la, f.in[which],
                                                            nullptr, f.b, f.k[ks], e.blk * 2 * C::PAIRS);
    });
}

template <typename T>
int batch_row_blocks(int kind, int lg, int rq, int ny) {
    if (kind == 1) {
        // column pass: blocks of G items; ny here is the item count
        int g = 1;
        dispatch_log2<1, Range<T>::COL>(lg, [&](auto lgc) {
            constexpr int LG = decltype(lgc)::value;
            if constexpr (col_rq_ok<LG, 3>()) {
                g = (rq & 7) == 3 ? ColGrp<T, LG, 3>::G : ColGrp<T, LG, 4>::G;
            } else {
                g = ColGrp<T, LG, 4>::G;
            }
        });
        return (ny + g - 1) / g;
    }
    int pairs = 1;
    dispatch_log2<1, Range<T>::ROW>(lg, [&](auto lgc) {
        constexpr int LG = decltype(lgc)::value;
        if (kind == 3) {
            pairs = rq == 3 ? BatchFusedCfg<T, LG, 3>::PAIRS : BatchFusedCfg<T, LG, 4>::PAIRS;
        } else {
            pairs = rq == 3 ? RowCfg<T, LG, 3>::PAIRS : RowCfg<T, LG, 4>::PAIRS;
        }
    });
    return (ny + 2 * pairs - 1) / (2 * pairs);
}

// Calls fn(KIND, LO, HI, RQ as integral constants) for the merged class of
// (kind, lg, rq).
template <typename T, typename Fn>
static void batch_class_dispatch(int kind, int lg, int rq, Fn &&fn) {
    auto with_kind = [&](auto kc) {
        constexpr int KIND = decltype(kc)::value;
        dispatch_log2<1, batch_max_lg<T, KIND>()>(lg, [&](auto lgc) {
            constexpr int LG = decltype(lgc)::value;
            auto with_rq = [&](auto rqc) {
                constexpr int RQ = decltype(rqc)::value;
                if constexpr (BatchCfg<T, KIND, LG, RQ>::OK) {
                    constexpr int LO = batch_run_lo<T, KIND, RQ, LG>(), HI = batch_run_hi<T, KIND, RQ, LG>();
                    fn(kc, std::integral_constant<int, LO>{}, std::integral_constant<int, HI>{}, rqc);
                } else {
                    throw Error(3, "batched FFT pass: no kernel for this length and radix");
                }
            };
            if ((rq & 7) == 3) with_rq(std::integral_constant<int, 3>{});
            else with_rq(std::integral_constant<int, 4>{});
        });
    };
    switch (kind) {
        case 0: with_kind(std::integral_constant<int, 0>{}); break;
        case 1: with_kind(std::integral_constant<int, 1>{}); break;
        case 2: with_kind(std::integral_constant<int, 2>{}); break;
        case 3: with_kind(std::integral_constant<int, 3>{}); break;
        default: throw Error(1, "batched FFT pass: unknown kind");
    }
}

template <typename T>
int batch_class_lg(int kind, int lg, int rq) {
    int lo = lg;
    batch_class_dispatch<T>(kind, lg, rq, [&](auto, auto loc, auto, auto) { lo = decltype(loc)::value; });
    return lo;
}

template <typename T>
void batch_allow_lds(int kind, int lg, int rq) {
    batch_class_dispatch<T>(kind, lg, rq, [&](auto kc, auto loc, auto hic, auto rqc) {
        constexpr int KIND = decltype(kc)::value, LO = decltype(loc)::value, HI = decltype(hic)::value,
                      RQ = decltype(rqc)::value;
        constexpr size_t lds = batch_run_lds<T, KIND, RQ, LO, HI>();
        if constexpr (KIND == 0) {
            allow_lds(kb_row_r2c<T, LO, HI, true, RQ>, lds);
            allow_lds(kb_row_r2c<T, LO, HI, false, RQ>, lds);
        } else if constexpr (KIND == 1) {
            allow_lds(kb_col<T, LO, HI, true, RQ>, lds);
            allow_lds(kb_col<T, LO, HI, false, RQ>, lds);
        } else if constexpr (KIND == 2) {
            allow_lds(kb_row_c2r<T, LO, HI, true, RQ>, lds);
            allow_lds(kb_row_c2r<T, LO, HI, false, RQ>, lds);
        } else {
            allow_lds(kb_row_c2r_local<T, LO, HI, true, RQ>, lds);
            allow_lds(kb_row_c2r_local<T, LO, HI, false, RQ>, lds);
        }
    });
}

template <typename T>
void launch_batch_fft(int kind, int lg, int rq, bool half, const BFilm<T> *F, const BEntry *tab, int nent,
                      int which, int ks, int stage1, hipStream_t s) {
    if (nent <= 0) return;
    const dim3 grid((unsigned)nent);
    batch_class_dispatch<T>(kind, lg, rq, [&](auto kc, auto loc, auto hic, auto rqc) {
        constexpr int KIND = decltype(kc)::value, LO = decltype(loc)::value, HI = decltype(hic)::value,
                      RQ = decltype(rqc)::value;
        constexpr size_t lds = batch_run_lds<T, KIND, RQ, LO, HI>();
        const dim3 nt((unsigned)BatchCfg<T, KIND, LO, RQ>::NT);
        if constexpr (KIND == 0) {
            if (half) hipLaunchKernelGGL((kb_row_r2c<T, LO, HI, true, RQ>), grid, nt, lds, s, F, tab, which, stage1);
            else hipLaunchKernelGGL((kb_row_r2c<T, LO, HI, false, RQ>), grid, nt, lds, s, F, tab, which, stage1);
        } else if constexpr (KIND == 1) {
            if (half) hipLaunchKernelGGL((kb_col<T, LO, HI, true, RQ>), grid, nt, lds, s, F, tab, stage1);
            else hipLaunchKernelGGL((kb_col<T, LO, HI, false, RQ>), grid, nt, lds, s, F, tab, stage1);
        } else if constexpr (KIND == 2) {
            if (half) hipLaunchKernelGGL((kb_row_c2r<T, LO, HI, true, RQ>), grid, nt, lds, s, F, tab, stage1);
            else hipLaunchKernelGGL((kb_row_c2r<T, LO, HI, false, RQ>), grid, nt, lds, s, F, tab, stage1);
        } else {
            if (half)
                hipLaunchKernelGGL((kb_row_c2r_local<T, LO, HI, true, RQ>), grid, nt, lds, s, F, tab, which, ks,
                                   stage1);
            else
                hipLaunchKernelGGL((kb_row_c2r_local<T, LO, HI, false, RQ>), grid, nt, lds, s, F, tab, which, ks,
                                   stage1);
        }
    });
    LLGMI_LAUNCH_CHECK();
}

template int batch_row_blocks<float>(int, int, int, int);
template int batch_row_blocks<double>(int, int, int, int);
template void batch_allow_lds<float>(int, int, int);
template int batch_class_lg<float>(int, int, int);
template void batch_allow_lds<double>(int, int, int);
template int batch_class_lg<double>(int, int, int);
template void launch_batch_fft<float>(int, int, int, bool, const BFilm<float> *, const BEntry *, int, int, int, int,
                                      hipStream_t);
template void launch_batch_fft<double>(int, int, int, bool, const BFilm<double> *, const BEntry *, int, int, int,
                                       int, hipStream_t);

template struct DemagFFT<float>;
template struct DemagFFT<double>;

}  // namespace llgmi

// Internal definitions shared by the HIP translation units of libllgmi.so.
// Target: gfx950 (MI355X), wave64.  Not part of the public ABI (include/llgmi.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace llgmi {

// ---------------------------------------------------------------------------
// Errors: internal code throws; the C ABI layer converts to status codes.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LLGMI_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::llgmi::Error(2, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LLGMI_LAUNCH_CHECK() LLGMI_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Complex value of precision T (plain struct: 8 or 16 bytes, naturally aligned)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) {
    return INV ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

// Four real kernel values at one k-space point: (Kxx, Kxy, Kyy, Kzz).
template <typename T>
struct alignas(4 * sizeof(T)) quad4 {
    T xx, xy, yy, zz;
};

// ---------------------------------------------------------------------------
// In-register DFTs of size 2, 4, 8 (forward: exp(-2 pi i qr/R); INV: conjugate)
// ---------------------------------------------------------------------------
template <typename T, bool INV>
__device__ __forceinline__ void dft2(cpx<T> *v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cpx<T> *v) {
    cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cpx<T> a2 = cadd(v[1], v[3]), a3 = mul_mi<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[1] = cadd(a1, a3);
    v[2] = csub(a0, a2);
    v[3] = csub(a1, a3);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cpx<T> *v) {
    const T h = T(0.70710678118654752440084436210485);
    // radix-2 split: a_r = v_r + v_{r+4} (even outputs), b_r = (v_r - v_{r+4}) w8^r (odd)
    cpx<T> a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        a[r] = cadd(v[r], v[r + 4]);
        b[r] = csub(v[r], v[r + 4]);
    }
    // w8 = exp(-i pi/4) forward; conj for inverse
    b[1] = INV ? cpx<T>{h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)}
               : cpx<T>{h * (b[1].x + b[1].y), h * (b[1].y - b[1].x)};
    b[2] = mul_mi<T, INV>(b[2]);
    b[3] = INV ? cpx<T>{-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y)}
               : cpx<T>{h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y)};
    dft4<T, INV>(a);
    dft4<T, INV>(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[2 * q] = a[q];
        v[2 * q + 1] = b[q];
    }
}

template <typename T, bool INV, int R>
__device__ __forceinline__ void dftR(cpx<T> *v) {
    if constexpr (R == 2) dft2<T, INV>(v);
    else if constexpr (R == 4) dft4<T, INV>(v);
    else dft8<T, INV>(v);
}

// ---------------------------------------------------------------------------
// One Stockham stage over `nseq` sequences of length L held in LDS at
// buf[s*sstride + i].  Stage with span Ns and radix R:
//   read   v_r = buf[j + r*L/R] * w^(r*k), k = j mod Ns, w = exp(-2 pi i/(Ns R))
//   DFT_R(v)
//   write  buf[(j/Ns)*Ns*R + k + r*Ns] = v_r
// All reads of a stage complete (barrier) before any write, so one buffer
// suffices; each thread keeps at most EMAX/R butterflies in registers.
// tw[t] = exp(-2 pi i t / L) (t < L), precomputed in FP64 and rounded to T.
// ---------------------------------------------------------------------------
template <typename T, bool INV, int R, int EMAX>
__device__ __forceinline__ void stockham_stage(cpx<T> *buf, int L, int log2L, int nseq,
                                               int sstride, int Ns, int log2Ns,
                                               const cpx<T> *__restrict__ tw) {
    constexpr int BPT = (EMAX + R - 1) / R;
    const int nt = blockDim.x;
    const int tid = threadIdx.x;
    const int log2R = R == 8 ? 3 : (R == 4 ? 2 : 1);
    const int log2nb1 = log2L - log2R;  // butterflies per sequence = L/R
    const int nb = nseq << log2nb1;
    const int twstep_log2 = log2L - log2Ns - log2R;  // L/(Ns R)
    cpx<T> v[BPT][R];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        int b = tid + q * nt;
        if (b < nb) {
            int s = b >> log2nb1;
            int j = b & ((1 << log2nb1) - 1);
            int k = j & (Ns - 1);
            const cpx<T> *src = buf + s * sstride + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[r << log2nb1];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cpx<T> w = tw[(r * k) << twstep_log2];
                    v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
                }
            }
            dftR<T, INV, R>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        int b = tid + q * nt;
        if (b < nb) {
            int s = b >> log2nb1;
            int j = b & ((1 << log2nb1) - 1);
            int k = j & (Ns - 1);
            int idxD = ((j >> log2Ns) << (log2Ns + log2R)) + k;
            cpx<T> *dst = buf + s * sstride + idxD;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r << log2Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of nseq sequences of length L = 2^log2L in LDS (unnormalised).
// Stages: radix 8 while >= 3 bits remain except that a remainder of 1 bit after
// radix-8 stages is folded into two radix-4 stages (e.g. 2^10 = 8*8*4*4).
// Caller guarantees nseq*L <= EMAX*blockDim.x and a barrier before the call.
template <typename T, bool INV, int EMAX>
__device__ void lds_fft(cpx<T> *buf, int L, int log2L, int nseq, int sstride,
                        const cpx<T> *__restrict__ tw) {
    int rem = log2L;
    int Ns = 1, log2Ns = 0;
    while (rem > 0) {
        if (rem >= 3 && rem != 4) {
            stockham_stage<T, INV, 8, EMAX>(buf, L, log2L, nseq, sstride, Ns, log2Ns, tw);
            rem -= 3; Ns <<= 3; log2Ns += 3;
        } else if (rem >= 2) {
            stockham_stage<T, INV, 4, EMAX>(buf, L, log2L, nseq, sstride, Ns, log2Ns, tw);
            rem -= 2; Ns <<= 2; log2Ns += 2;
        } else {
            stockham_stage<T, INV, 2, EMAX>(buf, L, log2L, nseq, sstride, Ns, log2Ns, tw);
            rem -= 1; Ns <<= 1; log2Ns += 1;
        }
    }
}

inline int ilog2(int64_t v) {
    int r = 0;
    while ((int64_t(1) << r) < v) ++r;
    return r;
}
inline int64_t next_pow2(int64_t v) {
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace llgmi

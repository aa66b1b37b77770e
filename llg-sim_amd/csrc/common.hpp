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

// Records msg as the library's last error (llgmi_last_error) and returns code.
int set_last_error(int code, const std::string &msg);
// "" when [p, p + bytes) lies in one device allocation, else the reason.
std::string check_device_range(const void *p, size_t bytes, const char *what);

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

// XCD-aware block order.  Workgroups are dealt round-robin over the eight XCDs
// of an MI355X (block p runs on XCD p % 8), and each XCD has its own L2.  The
// map below gives each XCD a contiguous range of logical blocks, so blocks
// that share cache lines (neighbouring rows, row pairs of one line) are
// resident in the same L2.  It is a bijection on [0, n).
__device__ __forceinline__ int xcd_remap(int p, int n) {
    constexpr int NX = 8;
    const int q = n / NX, r = n % NX;
    const int xcd = p % NX, i = p / NX;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + i;
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

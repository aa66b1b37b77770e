// Register-resident Stockham FFTs for gfx950 (power-of-two lengths).
//
// One sequence of length L = 2^LOG2L is transformed by TPS = L / R0 threads,
// each holding R0 complex values in registers.  The stage radices are R0 for
// the first and the last stage and at most R0 in between; middle stages go
// through LDS.  Because the first and the last stage have the same radix and
// every thread owns exactly one butterfly in both, a thread's register view of
// the sequence is the same before the first stage and after the last one:
//
//     v[r] = x[t + r * TPS],  r < R0,  t = thread index within the sequence.
//
// So a transform can start from values a thread loaded itself from HBM and
// end with values it can store (or multiply and inverse-transform) directly:
// a forward transform, a pointwise product and an inverse transform need no
// reordering pass.  L <= 16 is one in-register DFT with no LDS at all.
//
// LDS layout of a sequence: element i at lds_pad(i) = i + i/16, which spreads
// the stride-R0 writes of the first stage over the banks.  Every access below
// is a per-thread base plus a compile-time offset: for the power-of-two
// strides used here lds_pad(a + b) = lds_pad(a) + lds_pad(b) whenever b is a
// multiple of 16, or a's position within its aligned 16-element group plus b
// stays inside that group (checked exhaustively for every plan), so the
// compiler emits one address per stage and immediate ds_read/ds_write offsets.
#pragma once

#include "common.hpp"

namespace llgmi {

__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 4); }
// Eight extra slots per sequence: padded power-of-two lengths are multiples of
// the 64 four-byte banks, so without the offset every sequence of a block
// starts on bank 0 and the row passes' column-wise reads and writes (one
// element from each of several sequences) collide.  An offset of 8 complex
// values (16 banks) puts four consecutive sequences on disjoint bank ranges.
__host__ __device__ constexpr int lds_seq_len(int L) { return L + (L >> 4) + 8; }

// In-register DFT16 as 4 x 4: X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[4 n1 + n2] W4^(n1 k1)
template <typename T, bool INV>
__device__ __forceinline__ void dft16(cpx<T> *v) {
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);  // cos, sin pi/8
    const T h = T(0.70710678118654752440);
    cpx<T> a[4][4];  // a[n2][k1]
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cpx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // twiddles W16^(n2 k1), W16 = exp(-i pi/8) (conjugate for INV)
    auto tw = [&](cpx<T> x, T c, T s) -> cpx<T> {
        // multiply by (c - i s) forward, (c + i s) inverse
        return INV ? cpx<T>{x.x * c - x.y * s, x.x * s + x.y * c}
                   : cpx<T>{x.x * c + x.y * s, x.y * c - x.x * s};
    };
    a[1][1] = tw(a[1][1], c1, s1);                 // W^1
    a[1][2] = tw(a[1][2], h, h);                   // W^2
    a[1][3] = tw(a[1][3], s1, c1);                 // W^3
    a[2][1] = tw(a[2][1], h, h);                   // W^2
    a[2][2] = mul_mi<T, INV>(a[2][2]);             // W^4 = -i
    a[2][3] = tw(a[2][3], -h, h);                  // W^6
    a[3][1] = tw(a[3][1], s1, c1);                 // W^3
    a[3][2] = tw(a[3][2], -h, h);                  // W^6
    a[3][3] = tw(a[3][3], -c1, -s1);               // W^9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx<T> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

template <typename T, bool INV, int R>
__device__ __forceinline__ void dft_reg(cpx<T> *v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) dft2<T, INV>(v);
    else if constexpr (R == 4) dft4<T, INV>(v);
    else if constexpr (R == 8) dft8<T, INV>(v);
    else dft16<T, INV>(v);
}

// RQ: requested log2 of the first/last radix (0 = default); ignored unless
// both stages fit (2*RQ <= LOG2L).
template <int LOG2L, int RQ = 0>
struct RegPlan {
    static constexpr int L = 1 << LOG2L;
    static constexpr int DEF = LOG2L <= 4 ? LOG2L : (LOG2L >= 8 ? 4 : (LOG2L >= 6 ? 3 : 2));
    static constexpr int LOG2R0 = (RQ > 0 && 2 * RQ <= LOG2L) ? RQ : DEF;
    static constexpr int R0 = 1 << LOG2R0;
    static constexpr int TPS = L / R0;                                   // threads per sequence
    static constexpr int MID = LOG2L <= 4 ? 0 : LOG2L - 2 * LOG2R0;      // bits of the middle stages
    // radix (log2) of the stage starting with LOG2NS bits done
    static constexpr int log2r(int log2ns) {
        if (log2ns == 0) return LOG2R0;
        int rem = LOG2L - log2ns;
        if (rem == LOG2R0) return LOG2R0;                                // last stage
        int mid_rem = rem - LOG2R0;
        return mid_rem >= LOG2R0 ? LOG2R0 : mid_rem;
    }
};

// Twiddles of stage (Ns, R) for butterfly class k: w^r with w = tw[k L/(Ns R)],
// r < R.  Only w, w^4 and w^8 are gathered from the table, before the stage's
// barrier so that the load latency overlaps the wait; the other powers are
// products of at most three table values (a few ulp in FP32).
template <typename T, int R>
struct TwBase {
    cpx<T> w1, w4, w8;
};
template <typename T, int R>
__device__ __forceinline__ TwBase<T, R> twiddle_load(const cpx<T> *__restrict__ tw, int base) {
    TwBase<T, R> b;
    b.w1 = tw[base];
    if constexpr (R >= 8) b.w4 = tw[4 * base];
    if constexpr (R == 16) b.w8 = tw[8 * base];
    return b;
}
// All powers w^r, r = 1..R-1, from the table values (formed once per stage
// and thread, shared by the thread's butterflies of that stage).
template <typename T, int R>
__device__ __forceinline__ void twiddle_powers(const TwBase<T, R> &b, cpx<T> *w) {
    w[1] = b.w1;
    if constexpr (R > 2) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[2], w[1]);
    }
    if constexpr (R >= 8) {
        w[4] = b.w4;
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R == 16) {
        w[8] = b.w8;
#pragma unroll
        for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
    }
}
template <typename T, bool INV, int R>
__device__ __forceinline__ void twiddle_apply(cpx<T> *u, const cpx<T> *w) {
#pragma unroll
    for (int r = 1; r < R; ++r) u[r] = INV ? cmulc(u[r], w[r]) : cmul(u[r], w[r]);
}

// Middle/last stages.  Called after the first stage wrote its outputs to LDS.
// If this is the last stage its outputs stay in v.  The butterfly class k of
// a thread is the same for all its butterflies of a stage (NS divides TPS),
// so one set of twiddles serves them all.
template <typename T, bool INV, int LOG2L, int RQ, int LOG2NS>
__device__ __forceinline__ void reg_fft_rest(cpx<T> *v, cpx<T> *seq, int t,
                                             const cpx<T> *__restrict__ tw) {
    using P = RegPlan<LOG2L, RQ>;
    if constexpr (LOG2NS < LOG2L) {
        constexpr int LOG2R = P::log2r(LOG2NS);
        constexpr int R = 1 << LOG2R;
        constexpr int NS = 1 << LOG2NS;
        constexpr int L = P::L, TPS = P::TPS, R0 = P::R0;
        constexpr int BPT = R0 / R;                  // butterflies per thread
        constexpr int NB = L / R;                    // butterflies per sequence
        constexpr int TWS = LOG2L - LOG2NS - LOG2R;
        constexpr bool LAST = LOG2NS + LOG2R == LOG2L;
        static_assert(TPS % NS == 0, "butterfly class must not depend on the butterfly");
        const TwBase<T, R> twb = twiddle_load<T, R>(tw, (t & (NS - 1)) << TWS);
        __syncthreads();  // previous stage's LDS writes are visible
        const cpx<T> *rb = seq + lds_pad(t);
        cpx<T> w[R];
        twiddle_powers<T, R>(twb, w);
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            cpx<T> *u = v + q * R;
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = rb[lds_pad(q * TPS) + lds_pad(r * NB)];
            twiddle_apply<T, INV, R>(u, w);
            dft_reg<T, INV, R>(u);
        }
        if constexpr (!LAST) {
            __syncthreads();  // all reads of this stage done before it overwrites
#pragma unroll
            for (int q = 0; q < BPT; ++q) {
                const int j = t + q * TPS;
                const int k = j & (NS - 1);
                const int idxD = ((j >> LOG2NS) << (LOG2NS + LOG2R)) + k;
                cpx<T> *wb = seq + lds_pad(idxD);
#pragma unroll
                for (int r = 0; r < R; ++r) wb[lds_pad(r * NS)] = v[q * R + r];
            }
            reg_fft_rest<T, INV, LOG2L, RQ, LOG2NS + LOG2R>(v, seq, t, tw);
        }
        // LAST: BPT == 1, j == t, outputs v[r] = X[t + r*TPS]
    }
}

// Transform of one sequence held as v[r] = x[t + r*TPS] by its TPS threads,
// using `seq` (lds_seq_len(L) slots) as exchange space.  On return
// v[r] = X[t + r*TPS].  All threads of the block must call it (it contains
// barriers); the caller must not touch `seq` between the call and a barrier
// that follows it.  The first LDS write happens after a leading barrier, so
// `seq` may still be read by the caller before the call.
template <typename T, bool INV, int LOG2L, int RQ = 0>
__device__ __forceinline__ void reg_fft(cpx<T> *v, cpx<T> *seq, int t,
                                        const cpx<T> *__restrict__ tw) {
    using P = RegPlan<LOG2L, RQ>;
    dft_reg<T, INV, P::R0>(v);
    if constexpr (LOG2L > P::LOG2R0) {
        __syncthreads();  // the caller may have been reading seq
        cpx<T> *wb = seq + lds_pad(t * P::R0);
#pragma unroll
        for (int r = 0; r < P::R0; ++r) wb[lds_pad(r)] = v[r];
        reg_fft_rest<T, INV, LOG2L, RQ, P::LOG2R0>(v, seq, t, tw);
    }
}

// ---------------------------------------------------------------------------
// Transform-length codes: c < kF3 stands for 2^c points, c >= kF3 for
// 3 * 2^(c - kF3) points (M = 2^(c - kF3)).  A length 3M is three register
// transforms of length M over the interleaved subsequences x[3m + g] and a
// radix-3 combination (decimation in time forward, in frequency inverse):
//   X[k + qM] = sum_g W^(g k) w3^(g q) Y_g[k],  W = exp(-2 pi i / 3M),
//   w3 = exp(-2 pi i / 3), Y_g = FFT_M(x[3m + g]).
// A thread then holds NV = 3 R0 values: in the frequency domain v[r'] =
// X[t + r' TPS] (as for a power of two), in the time domain v[r'] =
// x[tindex(t, r')].  Twiddle tables of a 3M code hold the M-point table and
// then the 3M-point table.
// ---------------------------------------------------------------------------
constexpr int kF3 = 32;
__host__ __device__ constexpr bool code_f3(int c) { return c >= kF3; }
__host__ __device__ constexpr int code_log2m(int c) { return c >= kF3 ? c - kF3 : c; }
__host__ __device__ constexpr int code_len(int c) { return (c >= kF3 ? 3 : 1) << code_log2m(c); }
// entries of a code's twiddle table
__host__ __device__ constexpr int code_tw_len(int c) { return c >= kF3 ? 4 << code_log2m(c) : 1 << c; }

template <int LC, int RQ = 0>
struct Plan {
    static constexpr bool F3 = code_f3(LC);
    static constexpr int LOG2M = code_log2m(LC);
    using P = RegPlan<LOG2M, RQ>;
    static constexpr int M = P::L, L = code_len(LC), R0 = P::R0, TPS = P::TPS;
    static constexpr int NV = F3 ? 3 * R0 : R0;   // values per thread
    // time-domain index of value rp of thread t (its frequency index is t + rp * TPS)
    __host__ __device__ static constexpr int tindex(int t, int rp) {
        return F3 ? 3 * (t + (rp % R0) * TPS) + rp / R0 : t + rp * TPS;
    }
    // whether that time-domain index lies below L / 2
    __host__ __device__ static constexpr bool lower(int rp) { return (rp % R0) < R0 / 2; }
};

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cpx<T> &a, cpx<T> &b, cpx<T> &c) {
    const T h = T(0.86602540378443864676);   // sin(2 pi / 3)
    const cpx<T> s = cadd(b, c), d = csub(b, c);
    const cpx<T> m = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
    const cpx<T> ihd = {-h * d.y, h * d.x};   // i h (b - c)
    a = cadd(a, s);
    if (INV) {
        b = cadd(m, ihd);
        c = csub(m, ihd);
    } else {
        b = csub(m, ihd);
        c = cadd(m, ihd);
    }
}

// Transform of length code LC: time domain -> frequency domain (forward) or
// back (INV), with the register views above.  `seq` holds lds_seq_len(M)
// slots for the sub-transforms (the caller's own exchanges need
// lds_seq_len(L)); the barrier rules of reg_fft apply.
template <typename T, bool INV, int LC, int RQ = 0>
__device__ __forceinline__ void plan_fft(cpx<T> *v, cpx<T> *seq, int t, const cpx<T> *__restrict__ tw) {
    using PL = Plan<LC, RQ>;
    if constexpr (!PL::F3) {
        reg_fft<T, INV, LC, RQ>(v, seq, t, tw);
    } else {
        constexpr int R0 = PL::R0, TPS = PL::TPS, M = PL::M;
        const cpx<T> *__restrict__ tw3 = tw + M;
        if constexpr (!INV) {
#pragma unroll
            for (int g = 0; g < 3; ++g) reg_fft<T, false, PL::LOG2M, RQ>(v + g * R0, seq, t, tw);
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int k = t + r * TPS;
                cpx<T> a = v[r], b = cmul(v[R0 + r], tw3[k]), c = cmul(v[2 * R0 + r], tw3[2 * k]);
                dft3<T, false>(a, b, c);
                v[r] = a;
                v[R0 + r] = b;
                v[2 * R0 + r] = c;
            }
        } else {
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int k = t + r * TPS;
                cpx<T> a = v[r], b = v[R0 + r], c = v[2 * R0 + r];
                dft3<T, true>(a, b, c);
                v[r] = a;
                v[R0 + r] = cmulc(b, tw3[k]);
                v[2 * R0 + r] = cmulc(c, tw3[2 * k]);
            }
#pragma unroll
            for (int g = 0; g < 3; ++g) reg_fft<T, true, PL::LOG2M, RQ>(v + g * R0, seq, t, tw);
        }
    }
}

// Compile-time dispatch over log2 lengths LO..HI.
template <int LO, int HI, typename F>
inline void dispatch_log2(int lg, F &&f) {
    if constexpr (LO <= HI) {
        if (lg == LO) {
            f(std::integral_constant<int, LO>{});
            return;
        }
        dispatch_log2<LO + 1, HI>(lg, static_cast<F &&>(f));
    } else {
        throw Error(3, "FFT length 2^" + std::to_string(lg) + " is outside the built range");
    }
}

}  // namespace llgmi

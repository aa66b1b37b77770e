// Compile-time specialised LDS FFTs (power-of-two lengths) for gfx950.
//
// A block transforms NSEQ sequences of length L = 2^LOG2L held in LDS, with NT
// threads.  Stockham autosort, radix 8 (radix 4 for the last 2 or 4 bits,
// radix 2 only for L = 2).  Every stage is unrolled at compile time: each
// thread holds exactly ceil(NSEQ*L/R / NT) butterflies in registers between
// the read and write halves of a stage, so nothing spills to scratch.
//
// LDS index padding: element i of a sequence is stored at i + (i >> 4) (one
// pad slot per 16 complex values), which breaks the power-of-two strides of
// the Stockham writes across banks.
#pragma once

#include "common.hpp"

namespace llgmi {

__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_seq_len(int L) { return L + (L >> 4); }

template <int LOG2L>
struct FftStage {
    // radix of the stage that starts with LOG2NS bits already done
    static constexpr int log2r(int log2ns) {
        int rem = LOG2L - log2ns;
        return (rem >= 3 && rem != 4) ? 3 : (rem >= 2 ? 2 : 1);
    }
};

template <typename T, bool INV, int LOG2L, int NSEQ, int NT, int LOG2NS>
__device__ __forceinline__ void fft_stage_rec(cpx<T> *buf, const cpx<T> *__restrict__ tw) {
    if constexpr (LOG2NS < LOG2L) {
        constexpr int LOG2R = FftStage<LOG2L>::log2r(LOG2NS);
        constexpr int R = 1 << LOG2R;
        constexpr int L = 1 << LOG2L;
        constexpr int NS = 1 << LOG2NS;
        constexpr int NB = L / R;              // butterflies per sequence
        constexpr int TOT = NB * NSEQ;
        constexpr int BPT = (TOT + NT - 1) / NT;
        constexpr int SLEN = lds_seq_len(L);
        constexpr int TWS = LOG2L - LOG2NS - LOG2R;  // twiddle step log2: L/(Ns R)
        const int tid = threadIdx.x;
        cpx<T> v[BPT][R];
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int b = tid + q * NT;
            if ((TOT % NT == 0) || b < TOT) {
                const int s = b / NB;
                const int j = b % NB;
                const int k = j & (NS - 1);
                const cpx<T> *src = buf + s * SLEN;
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = src[lds_pad(j + r * NB)];
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        const cpx<T> w = tw[(r * k) << TWS];
                        v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
                    }
                }
                dftR<T, INV, R>(v[q]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int b = tid + q * NT;
            if ((TOT % NT == 0) || b < TOT) {
                const int s = b / NB;
                const int j = b % NB;
                const int k = j & (NS - 1);
                const int idxD = ((j >> LOG2NS) << (LOG2NS + LOG2R)) + k;
                cpx<T> *dst = buf + s * SLEN;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[lds_pad(idxD + r * NS)] = v[q][r];
            }
        }
        __syncthreads();
        fft_stage_rec<T, INV, LOG2L, NSEQ, NT, LOG2NS + LOG2R>(buf, tw);
    }
}

// Unnormalised transform of NSEQ padded sequences (stride lds_seq_len(L)).
// Requires a barrier between the last writes to buf and the call; ends with one.
template <typename T, bool INV, int LOG2L, int NSEQ, int NT>
__device__ __forceinline__ void lds_fft_ct(cpx<T> *buf, const cpx<T> *__restrict__ tw) {
    fft_stage_rec<T, INV, LOG2L, NSEQ, NT, 0>(buf, tw);
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

// unet.hip -- learned initial guess for the multigrid Poisson demag.
//
// A 2-D U-Net maps (m_x, m_y, m_z, film mask) on the plane of the MG box to
// the box potential, one output channel per z plane of the box.  Weights and
// activations are bf16 on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32
// accumulation); gradients, master weights and Adam moments are FP32.  All of
// training runs here: forward, backward (data and weight gradients as
// implicit GEMMs), loss and optimiser.  The multigrid solver still decides
// convergence: the prediction only replaces the initial phi, and the solve
// iterates to the same residual (the reference has no counterpart; its solver starts from
// zero or from the previous solve, demag_poisson_mg.rs:1947-2057).
//
// Layout: activations channels-last, pixel p = (b*H + y)*W + x for batch b,
// row y, column x; a tensor slice is (pointer, channel stride, channel
// offset).  Convolution weights [tap][c_out][c_in] with tap = (dy+1)*3+(dx+1),
// zero padding at the image border.
//
// Architecture (levels l = 0..3, c_l = base * 2^l):
//   enc_l: conv3x3(c_{l-1} or c_in -> c_l) + ReLU, conv3x3(c_l -> c_l) + ReLU,
//          2x2 average pool between levels;
//   dec_l (l = 2..0): nearest 2x upsample of the level below, concatenated
//          [up, skip] with enc_l's output, then two conv3x3 + ReLU to c_l;
//   head:  conv3x3(c_0 -> c_out), linear, FP32 output.
#include "../../include/llgmi.h"
#include "common.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

namespace llgmi {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLevels = 4;
constexpr int kNT = 256;
// levels with at least this many pixels use 64-pixel waves (4 row tiles)
constexpr int64_t kWideTileMinPixels = 65536;
// dynamic LDS of a convolution block (the per-tap weight slice)
constexpr size_t kConvLdsMax = 64 * 1024;

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }


// ---------------------------------------------------------------------------
// 3x3 convolution as an implicit GEMM: rows = pixels, columns = output
// channels, reduction = 9 taps x input channels.  A block is 4 waves; a wave
// owns 16*MT pixels (MT 16-row MFMA tiles) x 16*NT channels; the 4 waves
// share the block's channels.  Per tap, the block's weight slice
// (16*NT rows of cin, rows padded by 8 elements against bank conflicts) is
// staged in LDS once and read from there by all waves; the pixel operands are
// 8-byte loads of 4 consecutive channels from global memory.  Also the data
// gradient of a convolution: the same product of the output gradient with
// the tap-flipped, transposed weights.
// Requires: cin % 16 == 0, cout % (16*NT) == 0, xs, xo, ys, yo multiples of 4,
// dynamic LDS of 16*NT*(cin+8)*2 bytes.
// ---------------------------------------------------------------------------
template <int MT, int NT, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ x, int xs, int xo, int cin,
                                                 const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                 void *__restrict__ y, int ys, int yo, int cout, int B, int H,
                                                 int W) {
    extern __shared__ __align__(16) unsigned char conv_smem[];
    uint16_t *wsl = reinterpret_cast<uint16_t *>(conv_smem);
    const int ws_row = cin + 8;   // LDS row stride of the weight slice (elements)
    const int P = B * H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    const int p_base = (blockIdx.x * 4 + wave) * (16 * MT);
    const int co_base = blockIdx.y * (16 * NT);
    int yl[MT], xl[MT];
    bool pv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int p = p_base + mt * 16 + li;
        pv[mt] = p < P;
        xl[mt] = p % W;
        yl[mt] = (p / W) % H;
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const uint16_t *xa[MT];
        bool qv[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int yy = yl[mt] + dy, xx = xl[mt] + dx;
            qv[mt] = pv[mt] && yy >= 0 && yy < H && xx >= 0 && xx < W;
            const int64_t q = (int64_t)(p_base + mt * 16 + li) + dy * W + dx;
            xa[mt] = x + (qv[mt] ? q * xs + xo + lk : 0);
        }
        // stage w[tap][co_base .. co_base + 16 NT)[0 .. cin) (contiguous) in LDS, 16 bytes at a time
        __syncthreads();   // the previous tap's reads are done
        {
            const uint4 *src = reinterpret_cast<const uint4 *>(w + ((int64_t)tap * cout + co_base) * cin);
            const int per_row = cin / 8, total = 16 * NT * per_row;
            for (int e = threadIdx.x; e < total; e += 256) {
                const int row = e / per_row, c8 = e - row * per_row;
                *reinterpret_cast<uint4 *>(wsl + row * ws_row + c8 * 8) = src[e];
            }
        }
        __syncthreads();
        const uint16_t *wt = wsl + li * ws_row + lk;
        for (int ci0 = 0; ci0 < cin; ci0 += 16) {
            bf16x4 a[MT], b[NT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[mt] = qv[mt] ? *(const bf16x4 *)(xa[mt] + ci0) : bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = *(const bf16x4 *)(wt + nt * 16 * ws_row + ci0);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
        }
    }
    // lane holds rows 4*(lane/16)+r, column lane%16 of every 16x16 tile
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = co_base + nt * 16 + li;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = p_base + mt * 16 + lk + r;
                if (p >= P) continue;
                float v = acc[mt][nt][r] + bv;
                if (RELU) v = fmaxf(v, 0.f);
                const int64_t o = (int64_t)p * ys + yo + co;
                if (OUT_F32) ((float *)y)[o] = v;
                else ((uint16_t *)y)[o] = f2bf(v);
            }
        }
}

// 3x3 convolution on spatial tiles (gfx950's K = 32 bf16 matrix-core
// product, v_mfma_f32_16x16x32_bf16): a block computes a TH x 32 pixel tile
// of one image for 16*NT output channels.  Input channels go in chunks of
// 32: the chunk's input tile with its one-pixel halo ((TH+2) x 34 pixels,
// zero outside the image) and the chunk's weights of all nine taps are staged
// in LDS once, then every tap's product reads both from LDS, so each input
// pixel is read from memory once per chunk instead of once per tap.  Wave w
// owns rows w*TH/4 .. of the tile (16-pixel half rows as MFMA row tiles) and
// all 16*NT channels.  The same product as k_conv3x3 (also used for the data
// gradient); it needs the pixel slices 16-byte aligned (xs, xo multiples of
// 8).  LDS rows are padded by 8 elements against bank conflicts.
constexpr int kTileW = 32, kTileKC = 32, kTileRow = kTileKC + 8;
template <int TH, int NT>
constexpr size_t conv_tile_lds() {
    return sizeof(uint16_t) * (size_t)kTileRow * ((TH + 2) * (kTileW + 2) + 9 * 16 * NT);
}
template <int TH, int NT, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3_tile(const uint16_t *__restrict__ x, int xs, int xo, int cin,
                                                      const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                      void *__restrict__ y, int ys, int yo, int cout, int B, int H,
                                                      int W) {
    constexpr int RW = TH / 4;           // tile rows per wave
    constexpr int MT = 2 * RW;           // 16-pixel row tiles per wave
    constexpr int CO = 16 * NT;          // output channels of the block
    constexpr int TPW = kTileW + 2;      // tile width with halo
    constexpr int NIN = (TH + 2) * TPW;  // staged input pixels
    extern __shared__ __align__(16) unsigned char tile_smem[];
    uint16_t *xin = reinterpret_cast<uint16_t *>(tile_smem);   // [NIN][kTileRow]
    uint16_t *wsl = xin + NIN * kTileRow;                      // [9][CO][kTileRow]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kg = (lane >> 4) * 8;
    const int ntx = (W + kTileW - 1) / kTileW, nty = (H + TH - 1) / TH;
    const int tx = blockIdx.x % ntx, ty = (blockIdx.x / ntx) % nty, b = blockIdx.x / (ntx * nty);
    const int x0 = tx * kTileW, y0 = ty * TH;
    const int co_base = blockIdx.y * CO;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    // A chunk's input tile and weights go through registers: the next chunk's
    // loads are issued before the current chunk's products, so they are in
    // flight while the matrix cores work.
    constexpr int NI = (NIN * (kTileKC / 8) + 255) / 256, NW = (9 * CO * (kTileKC / 8) + 255) / 256;
    uint4 ri[NI], rw[NW];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int e = threadIdx.x + k * 256;
            const int pix = e >> 2, c = c0 + (e & 3) * 8;
            const int yy = y0 + pix / TPW - 1, xx = x0 + pix % TPW - 1;
            ri[k] = uint4{0u, 0u, 0u, 0u};
            if (e < NIN * 4 && yy >= 0 && yy < H && xx >= 0 && xx < W && c < cin)
                ri[k] = *reinterpret_cast<const uint4 *>(x + ((int64_t)(b * H + yy) * W + xx) * xs + xo + c);
        }
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int e = threadIdx.x + k * 256;
            const int row = e >> 2, c = c0 + (e & 3) * 8;   // row = tap * CO + co
            const int tap = row / CO, co = row - tap * CO;
            rw[k] = uint4{0u, 0u, 0u, 0u};
            if (e < 9 * CO * 4 && c < cin)
                rw[k] = *reinterpret_cast<const uint4 *>(w + ((int64_t)tap * cout + co_base + co) * cin + c);
        }
    };
    fetch(0);
    for (int c0 = 0; c0 < cin; c0 += kTileKC) {
        __syncthreads();   // the previous chunk's reads are done
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int e = threadIdx.x + k * 256;
            if (e < NIN * 4) *reinterpret_cast<uint4 *>(xin + (e >> 2) * kTileRow + (e & 3) * 8) = ri[k];
        }
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int e = threadIdx.x + k * 256;
            if (e < 9 * CO * 4) *reinterpret_cast<uint4 *>(wsl + (e >> 2) * kTileRow + (e & 3) * 8) = rw[k];
        }
        __syncthreads();
        if (c0 + kTileKC < cin) fetch(c0 + kTileKC);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;   // offsets into the haloed tile
            bf16x8 a[MT], bw[NT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int r = wave * RW + (mt >> 1), col = (mt & 1) * 16 + li;
                a[mt] = *reinterpret_cast<const bf16x8 *>(xin + ((r + dy) * TPW + col + dx) * kTileRow + kg);
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bw[nt] = *reinterpret_cast<const bf16x8 *>(wsl + (tap * CO + nt * 16 + li) * kTileRow + kg);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], bw[nt], acc[mt][nt], 0, 0, 0);
        }
    }
    // lane holds pixels 4*(lane/16)+r of each 16-pixel row tile, channel lane%16
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int yy = y0 + wave * RW + (mt >> 1);
        if (yy >= H) continue;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = co_base + nt * 16 + li;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int xx = x0 + (mt & 1) * 16 + (lane >> 4) * 4 + r;
                if (xx >= W) continue;
                float v = acc[mt][nt][r] + bv;
                if (RELU) v = fmaxf(v, 0.f);
                const int64_t o = ((int64_t)(b * H + yy) * W + xx) * ys + yo + co;
                if (OUT_F32) ((float *)y)[o] = v;
                else ((uint16_t *)y)[o] = f2bf(v);
            }
        }
    }
}

// Weight gradient: dw[tap][co][ci] += sum_p dz[p][co] * x[p + off(tap)][ci].
// Block = 4 waves over one tap and one (16*MT) x (16*NT) tile and a pixel
// chunk, taken kSlice pixels at a time: the block loads 4-pixel x 4-channel
// blocks of dz and of the shifted x (four 8-byte loads each), transposes
// them in registers into LDS tiles [channel][pixel], and each wave multiplies
// its quarter of the slice in 16-pixel MFMA steps (8-byte LDS reads of 4
// consecutive pixels).
// The waves' partial tiles are summed through LDS and added to dw with FP32
// atomics.
template <int MT, int NT>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t *__restrict__ dz, int cout,
                                                       const uint16_t *__restrict__ x, int xs, int xo, int cin,
                                                       float *__restrict__ dw, int B, int H, int W, int chunk) {
    constexpr int kSlice = 64;          // pixels per LDS fill
    constexpr int kRow = kSlice + 4;    // LDS row stride (elements) of the transposed tiles
    __shared__ __align__(16) uint16_t sa[16 * MT][kRow];   // dz^T: [co][pixel]
    __shared__ __align__(16) uint16_t sb[16 * NT][kRow];   // x^T:  [ci][pixel]
    __shared__ f32x4 red[4][MT * NT][64];
    const int P = B * H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    const int tap = blockIdx.z, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int co_tiles = cout / (16 * MT);
    const int co_base = (blockIdx.y % co_tiles) * 16 * MT;
    const int ci_base = (blockIdx.y / co_tiles) * 16 * NT;
    const int p_begin = blockIdx.x * chunk;
    const int p_end = min(P, p_begin + chunk);
    constexpr int nA = kSlice / 4 * 4 * MT, nB = kSlice / 4 * 4 * NT;   // (pixel quad, channel quad) items
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int p0 = p_begin; p0 < p_end; p0 += kSlice) {
        __syncthreads();   // the previous slice's reads are done
        for (int e = threadIdx.x; e < nA + nB; e += 256) {
            const bool isA = e < nA;
            const int f = isA ? e : e - nA;
            const int cq = f % (isA ? 4 * MT : 4 * NT), pq = f / (isA ? 4 * MT : 4 * NT);
            bf16x4 v[4];
            const int pb = p0 + pq * 4;
            int xc = pb % W, yc = (pb / W) % H;   // column and row of pixel pb, advanced below
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = pb + r;
                v[r] = bf16x4{0, 0, 0, 0};
                if (p < p_end) {
                    if (isA) {
                        v[r] = *(const bf16x4 *)(dz + (int64_t)p * cout + co_base + cq * 4);
                    } else {
                        const int xx = xc + dx, yy = yc + dy;
                        if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                            const int64_t q = (int64_t)p + dy * W + dx;
                            v[r] = *(const bf16x4 *)(x + q * xs + xo + ci_base + cq * 4);
                        }
                    }
                }
                if (++xc == W) {   // next pixel starts a new row (and perhaps a new batch image)
                    xc = 0;
                    if (++yc == H) yc = 0;
                }
            }
            uint16_t(*t)[kRow] = isA ? sa : sb;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                *(bf16x4 *)&t[cq * 4 + c][pq * 4] = bf16x4{v[0][c], v[1][c], v[2][c], v[3][c]};
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kSlice / 64; ++ks) {
            const int kp = wave * (kSlice / 4) + ks * 16 + lk;   // this lane's 4 consecutive pixels
            bf16x4 a[MT], b[NT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) a[mt] = *(const bf16x4 *)&sa[mt * 16 + li][kp];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = *(const bf16x4 *)&sb[nt * 16 + li][kp];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) red[wave][mt * NT + nt][lane] = acc[mt][nt];
    __syncthreads();
    // thread t sums entry (tile t/64, lane t%64) of the four waves
    for (int e = threadIdx.x; e < MT * NT * 64; e += 256) {
        const int t = e / 64, ln = e % 64;
        const f32x4 sum = red[0][t][ln] + red[1][t][ln] + red[2][t][ln] + red[3][t][ln];
        const int mt = t / NT, nt = t % NT;
        const int ci = ci_base + nt * 16 + (ln & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = co_base + mt * 16 + (ln >> 4) * 4 + r;
            atomicAdd(dw + ((int64_t)tap * cout + co) * cin + ci, sum[r]);
        }
    }
}

// db[c] += sum over the block's pixel chunk of dz[p][c] (cout <= 256): the
// 256 threads form G = 256 / cout groups of cout channel lanes; group g sums
// pixels p0 + g, p0 + g + G, ...; the groups are reduced through LDS.
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dz, int cout, int P, int chunk,
                                                   float *__restrict__ db) {
    __shared__ float part[256];
    const int G = 256 / cout;
    const int t = threadIdx.x, g = t / cout, c = t % cout;
    const int p0 = blockIdx.x * chunk, p1 = min(P, p0 + chunk);
    float s = 0.f;
    if (g < G)
        for (int p = p0 + g; p < p1; p += G) s += bf2f(dz[(int64_t)p * cout + c]);
    part[t] = s;
    __syncthreads();
    if (t < cout) {
        float r = 0.f;
        for (int k = 0; k < G; ++k) r += part[k * cout + t];
        atomicAdd(db + t, r);
    }
}

// k_bias_grad eight channels (16 bytes) per thread: cout a multiple of 8 and
// at most 256 (groups of 8 channels per thread, pixel lanes over the rest)
__global__ __launch_bounds__(256) void k_bias_grad_v8(const uint16_t *__restrict__ dz, int cout, int P, int chunk,
                                                      float *__restrict__ db) {
    __shared__ float part[256 * 8];
    const int C8 = cout / 8, G = 256 / C8;
    const int t = threadIdx.x, g = t / C8, c8 = t - g * C8;
    const int p0 = blockIdx.x * chunk, p1 = min(P, p0 + chunk);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (g < G)
        for (int p = p0 + g; p < p1; p += G) {
            const uint4 q = *reinterpret_cast<const uint4 *>(dz + (int64_t)p * cout + c8 * 8);
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += bf2f(h[k]);
        }
#pragma unroll
    for (int k = 0; k < 8; ++k) part[t * 8 + k] = g < G ? s[k] : 0.f;
    __syncthreads();
    if (t < cout) {
        const int cg = t / 8, ck = t - cg * 8;
        float r = 0.f;
        for (int k = 0; k < G; ++k) r += part[(k * C8 + cg) * 8 + ck];
        atomicAdd(db + t, r);
    }
}

// 2x2 average pool of a slice: y[b][Y][X][c] = mean of x[b][2Y+i][2X+j][xo+c]
__global__ void k_pool2(const uint16_t *__restrict__ x, int xs, int xo, int C, uint16_t *__restrict__ y, int B,
                        int H, int W) {
    const int Hc = H / 2, Wc = W / 2;
    const int64_t n = (int64_t)B * Hc * Wc * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t pc = i / C;
        const int X = (int)(pc % Wc), Y = (int)((pc / Wc) % Hc), b = (int)(pc / ((int64_t)Wc * Hc));
        const int64_t p00 = ((int64_t)b * H + 2 * Y) * W + 2 * X;
        float s = bf2f(x[p00 * xs + xo + c]) + bf2f(x[(p00 + 1) * xs + xo + c]) + bf2f(x[(p00 + W) * xs + xo + c]) +
                  bf2f(x[(p00 + W + 1) * xs + xo + c]);
        y[i] = f2bf(0.25f * s);
    }
}

// k_pool2 eight channels (16 bytes) at a time: C, xs, xo multiples of 8
__global__ void k_pool2_v8(const uint16_t *__restrict__ x, int xs, int xo, int C, uint16_t *__restrict__ y, int B,
                           int H, int W) {
    const int Hc = H / 2, Wc = W / 2, C8 = C / 8;
    const int64_t n = (int64_t)B * Hc * Wc * C8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C8) * 8;
        const int64_t pc = i / C8;
        const int X = (int)(pc % Wc), Y = (int)((pc / Wc) % Hc), b = (int)(pc / ((int64_t)Wc * Hc));
        const int64_t p00 = ((int64_t)b * H + 2 * Y) * W + 2 * X;
        const uint4 q[4] = {*reinterpret_cast<const uint4 *>(x + p00 * xs + xo + c),
                            *reinterpret_cast<const uint4 *>(x + (p00 + 1) * xs + xo + c),
                            *reinterpret_cast<const uint4 *>(x + (p00 + W) * xs + xo + c),
                            *reinterpret_cast<const uint4 *>(x + (p00 + W + 1) * xs + xo + c)};
        const uint16_t *h[4] = {reinterpret_cast<const uint16_t *>(&q[0]), reinterpret_cast<const uint16_t *>(&q[1]),
                                reinterpret_cast<const uint16_t *>(&q[2]), reinterpret_cast<const uint16_t *>(&q[3])};
        uint4 o;
        uint16_t *ho = reinterpret_cast<uint16_t *>(&o);
#pragma unroll
        for (int k = 0; k < 8; ++k) ho[k] = f2bf(0.25f * (bf2f(h[0][k]) + bf2f(h[1][k]) + bf2f(h[2][k]) + bf2f(h[3][k])));
        *reinterpret_cast<uint4 *>(y + pc * C + c) = o;
    }
}

// gradient of k_pool2: dx[fine][c] += 0.25 dy[coarse][c]   (dense C, FP32)
__global__ void k_pool2_bwd(const float *__restrict__ dy, int C, float *__restrict__ dx, int B, int H, int W) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % W), yy = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
        const int64_t pc = ((int64_t)b * (H / 2) + yy / 2) * (W / 2) + x / 2;
        dx[i] += 0.25f * dy[pc * C + c];
    }
}

// nearest 2x upsample into a slice: y[fine][yo+c] = x[coarse][c] (x dense C)
__global__ void k_up2(const uint16_t *__restrict__ x, int C, uint16_t *__restrict__ y, int ys, int yo, int B, int H,
                      int W) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int xx = (int)(p % W), yy = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
        const int64_t pc = ((int64_t)b * (H / 2) + yy / 2) * (W / 2) + xx / 2;
        y[p * ys + yo + c] = x[pc * C + c];
    }
}

// k_up2 eight channels (16 bytes) at a time: C, ys, yo multiples of 8
__global__ void k_up2_v8(const uint16_t *__restrict__ x, int C, uint16_t *__restrict__ y, int ys, int yo, int B,
                         int H, int W) {
    const int C8 = C / 8;
    const int64_t n = (int64_t)B * H * W * C8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C8) * 8;
        const int64_t p = i / C8;
        const int xx = (int)(p % W), yy = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
        const int64_t pc = ((int64_t)b * (H / 2) + yy / 2) * (W / 2) + xx / 2;
        *reinterpret_cast<uint4 *>(y + p * ys + yo + c) = *reinterpret_cast<const uint4 *>(x + pc * C + c);
    }
}

// gradient of k_up2: dx[coarse][c] = sum of the four dy[fine][yo+c]
__global__ void k_up2_bwd(const float *__restrict__ dy, int ys, int yo, int C, float *__restrict__ dx, int B, int H,
                          int W) {
    const int Hc = H / 2, Wc = W / 2;
    const int64_t n = (int64_t)B * Hc * Wc * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t pc = i / C;
        const int X = (int)(pc % Wc), Y = (int)((pc / Wc) % Hc), b = (int)(pc / ((int64_t)Wc * Hc));
        const int64_t p00 = ((int64_t)b * H + 2 * Y) * W + 2 * X;
        dx[i] = dy[p00 * ys + yo + c] + dy[(p00 + 1) * ys + yo + c] + dy[(p00 + W) * ys + yo + c] +
                dy[(p00 + W + 1) * ys + yo + c];
    }
}

// dst[p][c] = src[p][so+c]  (the skip part of a concatenation's gradient)
__global__ void k_slice_copy(const float *__restrict__ src, int ss, int so, int C, float *__restrict__ dst, int64_t P) {
    const int64_t n = P * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        dst[i] = src[(i / C) * ss + so + c];
    }
}

// dz = bf16(dy * [y > 0]) with y a slice (ReLU layers), or bf16(dy) (y null)
// k_relu_bwd eight channels per thread: C, ys, yo multiples of 8
__global__ void k_relu_bwd_v8(const float *__restrict__ dy, const uint16_t *__restrict__ y, int ys, int yo, int C,
                              uint16_t *__restrict__ dz, int64_t P) {
    const int C8 = C / 8;
    const int64_t n = P * C8;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = i / C8;
        const int c = (int)(i - p * C8) * 8;
        const float4 g0 = *reinterpret_cast<const float4 *>(dy + p * C + c);
        const float4 g1 = *reinterpret_cast<const float4 *>(dy + p * C + c + 4);
        float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
        if (y) {
            const uint4 q = *reinterpret_cast<const uint4 *>(y + p * ys + yo + c);
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&q);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (!(bf2f(h[k]) > 0.f)) g[k] = 0.f;
        }
        uint4 o;
        uint16_t *ho = reinterpret_cast<uint16_t *>(&o);
#pragma unroll
        for (int k = 0; k < 8; ++k) ho[k] = f2bf(g[k]);
        *reinterpret_cast<uint4 *>(dz + p * C + c) = o;
    }
}

__global__ void k_relu_bwd(const float *__restrict__ dy, const uint16_t *__restrict__ y, int ys, int yo, int C,
                           uint16_t *__restrict__ dz, int64_t P) {
    const int64_t n = P * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float g = dy[i];
        if (y) {
            const int c = (int)(i % C);
            if (!(bf2f(y[(i / C) * ys + yo + c]) > 0.f)) g = 0.f;
        }
        dz[i] = f2bf(g);
    }
}

// Mean squared error over channels < cv and pixels with x < wv, y < hv:
// loss += sum diff^2 (FP64 atomics per block), dpred = 2 diff * inv_count.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ pred, const uint16_t *__restrict__ tgt, int C,
                                             int cv, int wv, int hv, int B, int H, int W, double inv_count,
                                             float *__restrict__ dpred, double *__restrict__ loss) {
    __shared__ double part[4];
    const int64_t n = (int64_t)B * H * W * C;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        float g = 0.f;
        if (c < cv && x < wv && y < hv) {
            const double d = (double)pred[i] - (double)bf2f(tgt[i]);
            s += d * d;
            g = (float)(2.0 * d * inv_count);
        }
        dpred[i] = g;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) * inv_count);
}

// Adam (Kingma & Ba 2015) on FP32 master parameters; bc1 = 1 - b1^t, bc2 = 1 - b2^t
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// bf16 weights for the forward product (same layout) and for the data
// gradient: wb[tap'][ci][co] = w[8 - tap'][co][ci]
__global__ void k_w_cast(const float *__restrict__ w, int cout, int cin, uint16_t *__restrict__ wf,
                         uint16_t *__restrict__ wb) {
    const int64_t n = (int64_t)9 * cout * cin;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const int co = (int)((i / cin) % cout);
        const int tap = (int)(i / ((int64_t)cin * cout));
        const uint16_t h = f2bf(w[i]);
        wf[i] = h;
        wb[((int64_t)(8 - tap) * cin + ci) * cout + co] = h;
    }
}

// Input packing: x[b][y][x][0..3] = (m_x, m_y, m_z, |m|^2 > 1e-30) of film
// cell (x - offx, y - offy), zero outside the film; channels 4..15 zero.
template <typename T>
__global__ void k_pack_input(const T *__restrict__ m, int nx, int ny, int offx, int offy, uint16_t *__restrict__ xo,
                             int H, int W) {
    const int64_t n = (int64_t)H * W;
    const int64_t nc = (int64_t)nx * ny;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(p % W) - offx, y = (int)(p / W) - offy;
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (x >= 0 && x < nx && y >= 0 && y < ny) {
            const int64_t c = (int64_t)y * nx + x;
            const double mx = m[c], my = m[nc + c], mz = m[2 * nc + c];
            v[0] = (float)mx;
            v[1] = (float)my;
            v[2] = (float)mz;
            v[3] = (mx * mx + my * my + mz * mz) > 1e-30 ? 1.f : 0.f;
        }
        uint16_t *o = xo + p * 16;
        for (int c = 0; c < 16; ++c) o[c] = c < 4 ? f2bf(v[c]) : (uint16_t)0;
    }
}

// box potential phi[z][y][x] (px x py x pz) -> target t[y][x][z] / scale
__global__ void k_pack_phi(const double *__restrict__ phi, int px, int py, int pz, double inv_scale,
                           uint16_t *__restrict__ t, int C, int H, int W) {
    const int64_t n = (int64_t)H * W * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % W), y = (int)(p / W);
        float v = 0.f;
        if (c < pz && x < px && y < py) v = (float)(phi[((int64_t)c * py + y) * px + x] * inv_scale);
        t[i] = f2bf(v);
    }
}

// network output y[y][x][z] (FP32) -> box potential phi[z][y][x] * scale
__global__ void k_unpack_phi(const float *__restrict__ yv, int C, int H, int W, double scale, int accumulate,
                             double *__restrict__ phi, int px, int py, int pz) {
    const int64_t n = (int64_t)px * py * pz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % px), y = (int)((i / px) % py), z = (int)(i / ((int64_t)px * py));
        const double v = (double)yv[((int64_t)y * W + x) * C + z] * scale;
        phi[i] = accumulate ? phi[i] + v : v;
    }
}

inline int grid_of(int64_t n) { return (int)std::min<int64_t>((n + kNT - 1) / kNT, 16384); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side: the network, its buffers and one training step.
// ---------------------------------------------------------------------------
struct UNetConv {
    int cin = 0, cout = 0;
    bool relu = true;
    int64_t w_off = 0, b_off = 0;   // offsets in the FP32 parameter vector
    uint16_t *wf = nullptr, *wb = nullptr;
};

struct UNet {
    int B, H, W, cin_real, cout_real, cout, base, wv, hv;
    hipStream_t stream = nullptr;
    std::vector<UNetConv> conv;   // enc (2 per level), dec (2 per level 2..0), head
    int64_t nparam = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_wbf = nullptr;    // bf16 forward and transposed weights of every conv
    int64_t t = 0;                // Adam step count
    double *d_loss = nullptr, *h_loss = nullptr;
    std::vector<void *> bufs;
    // activations (bf16): per level
    uint16_t *ea[kLevels] = {}, *eb[kLevels] = {}, *pool[kLevels] = {}, *cat[kLevels] = {}, *da[kLevels] = {},
             *db[kLevels] = {}, *dz[kLevels] = {};
    float *out = nullptr;
    // gradients (FP32)
    float *g_out = nullptr, *g_a[kLevels] = {}, *g_b[kLevels] = {}, *g_cat[kLevels] = {}, *g_eb[kLevels] = {},
          *g_pool[kLevels] = {};

    int c(int l) const { return base << l; }
    int hl(int l) const { return H >> l; }
    int wl(int l) const { return W >> l; }
    int64_t pl(int l) const { return (int64_t)B * hl(l) * wl(l); }

    template <typename U>
    U *dalloc(int64_t count) {
        void *p = nullptr;
        LLGMI_HIP(hipMalloc(&p, sizeof(U) * (size_t)std::max<int64_t>(count, 1)));
        bufs.push_back(p);
        return (U *)p;
    }

    UNet(int B_, int H_, int W_, int cin_, int cout_, int base_, int wv_, int hv_, uint64_t seed)
        : B(B_), H(H_), W(W_), cin_real(cin_), cout_real(cout_), cout((cout_ + 15) / 16 * 16), base(base_),
          wv(wv_), hv(hv_) {
        try {
            LLGMI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            // layers
            auto add = [&](int ci, int co, bool relu) {
                UNetConv cv;
                cv.cin = ci;
                cv.cout = co;
                cv.relu = relu;
                cv.w_off = nparam;
                nparam += (int64_t)9 * ci * co;
                cv.b_off = nparam;
                nparam += co;
                conv.push_back(cv);
            };
            for (int l = 0; l < kLevels; ++l) {
                add(l == 0 ? 16 : c(l - 1), c(l), true);
                add(c(l), c(l), true);
            }
            for (int l = kLevels - 2; l >= 0; --l) {
                add(c(l + 1) + c(l), c(l), true);
                add(c(l), c(l), true);
            }
            add(c(0), cout, false);
            d_p = dalloc<float>(nparam);
            d_g = dalloc<float>(nparam);
            d_m = dalloc<float>(nparam);
            d_v = dalloc<float>(nparam);
            int64_t nw = 0;
            for (auto &cv : conv) nw += 2 * (int64_t)9 * cv.cin * cv.cout;
            d_wbf = dalloc<uint16_t>(nw);
            int64_t o = 0;
            for (auto &cv : conv) {
                cv.wf = d_wbf + o;
                o += (int64_t)9 * cv.cin * cv.cout;
                cv.wb = d_wbf + o;
                o += (int64_t)9 * cv.cin * cv.cout;
            }
            d_loss = dalloc<double>(1);
            LLGMI_HIP(hipHostMalloc((void **)&h_loss, sizeof(double), hipHostMallocDefault));
            // activations and gradients
            for (int l = 0; l < kLevels; ++l) {
                const int64_t P = pl(l);
                ea[l] = dalloc<uint16_t>(P * c(l));
                if (l == kLevels - 1) eb[l] = dalloc<uint16_t>(P * c(l));
                else {
                    cat[l] = dalloc<uint16_t>(P * (c(l + 1) + c(l)));
                    eb[l] = cat[l];   // enc_l's output lives in the skip part of cat[l]
                    da[l] = dalloc<uint16_t>(P * c(l));
                    db[l] = dalloc<uint16_t>(P * c(l));
                    g_cat[l] = dalloc<float>(P * (c(l + 1) + c(l)));
                }
                if (l > 0) {
                    pool[l] = dalloc<uint16_t>(P * c(l - 1));
                    g_pool[l] = dalloc<float>(P * c(l - 1));
                }
                const int zmax = l < kLevels - 1 ? c(l + 1) + c(l) : c(l);
                dz[l] = dalloc<uint16_t>(P * std::max(zmax, l == 0 ? cout : 0));
                g_a[l] = dalloc<float>(P * c(l));
                g_b[l] = dalloc<float>(P * c(l));
                g_eb[l] = dalloc<float>(P * c(l));
            }
            out = dalloc<float>(pl(0) * cout);
            g_out = dalloc<float>(pl(0) * cout);
            init(seed);
        } catch (...) {
            release();
            throw;
        }
    }
    ~UNet() { release(); }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : bufs) (void)hipFree(p);
        bufs.clear();
        if (h_loss) (void)hipHostFree(h_loss);
        h_loss = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // He-normal weights (std sqrt(2 / fan_in), fan_in over the real input
    // channels), the head scaled by 0.1; biases zero; padded channels zero.
    void init(uint64_t seed) {
        std::vector<float> h((size_t)nparam, 0.f);
        uint64_t s = seed ^ 0x9e3779b97f4a7c15ull;
        auto next = [&]() {   // xorshift64*
            s ^= s >> 12;
            s ^= s << 25;
            s ^= s >> 27;
            return (s * 0x2545f4914f6cdd1dull) >> 11;
        };
        auto uni = [&]() { return ((double)next() + 0.5) / 9007199254740992.0; };
        for (size_t k = 0; k < conv.size(); ++k) {
            const UNetConv &cv = conv[k];
            const bool first = k == 0, head = k + 1 == conv.size();
            const int fan = 9 * (first ? cin_real : cv.cin);
            const double sd = std::sqrt(2.0 / fan) * (head ? 0.1 : 1.0);
            for (int tap = 0; tap < 9; ++tap)
                for (int co = 0; co < cv.cout; ++co)
                    for (int ci = 0; ci < cv.cin; ++ci) {
                        const double u1 = uni(), u2 = uni();
                        const double gz = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                        const bool live = (!first || ci < cin_real) && (!head || co < cout_real);
                        h[(size_t)(cv.w_off + ((int64_t)tap * cv.cout + co) * cv.cin + ci)] = live ? (float)(sd * gz) : 0.f;
                    }
        }
        set_params(h.data());
    }

    void set_params(const float *hp) {
        LLGMI_HIP(hipMemcpyAsync(d_p, hp, sizeof(float) * (size_t)nparam, hipMemcpyHostToDevice, stream));
        LLGMI_HIP(hipMemsetAsync(d_m, 0, sizeof(float) * (size_t)nparam, stream));
        LLGMI_HIP(hipMemsetAsync(d_v, 0, sizeof(float) * (size_t)nparam, stream));
        t = 0;
        cast_weights();
        LLGMI_HIP(hipStreamSynchronize(stream));
    }
    void cast_weights() {
        for (auto &cv : conv) {
            const int64_t n = (int64_t)9 * cv.cin * cv.cout;
            hipLaunchKernelGGL(k_w_cast, dim3(grid_of(n)), dim3(kNT), 0, stream, d_p + cv.w_off, cv.cout, cv.cin,
                               cv.wf, cv.wb);
        }
        LLGMI_LAUNCH_CHECK();
    }

    // ---- launch helpers with the shape checks the kernels rely on ----
    template <bool RELU, bool OUT_F32>
    void conv_fwd(const uint16_t *x, int xs, int xo, int cin, const uint16_t *w, const float *bias, void *y, int ys,
                  int yo, int co, int l) {
        if (cin % 16 || xs % 4 || xo % 4 || ys % 4 || yo % 4 || xo + cin > xs || yo + co > ys || co % 16)
            throw Error(LLGMI_ERR_INVALID, "internal: convolution shape");
        const int64_t P = pl(l);
        if (xs % 8 == 0 && xo % 8 == 0 && tile_conv_enabled()) {
            // spatial tiles of 8 rows, or 4 where 8 would leave the GPU short of blocks
            const int nt = co % 64 == 0 ? 4 : co % 48 == 0 ? 3 : co % 32 == 0 ? 2 : 1;
            const int64_t tiles8 = (int64_t)B * ((hl(l) + 7) / 8) * ((wl(l) + kTileW - 1) / kTileW);
            const bool th8 = tiles8 * (co / (16 * nt)) >= 1024;
            auto go = [&](auto thc, auto ntc) {
                constexpr int TH = decltype(thc)::value, NT = decltype(ntc)::value;
                launch_conv_tile<TH, NT, RELU, OUT_F32>(x, xs, xo, cin, w, bias, y, ys, yo, co, l);
            };
            using I8 = std::integral_constant<int, 8>;
            using I4 = std::integral_constant<int, 4>;
            switch (nt) {
                case 4: th8 ? go(I8{}, std::integral_constant<int, 4>{}) : go(I4{}, std::integral_constant<int, 4>{}); break;
                case 3: th8 ? go(I8{}, std::integral_constant<int, 3>{}) : go(I4{}, std::integral_constant<int, 3>{}); break;
                case 2: th8 ? go(I8{}, std::integral_constant<int, 2>{}) : go(I4{}, std::integral_constant<int, 2>{}); break;
                default: th8 ? go(I8{}, std::integral_constant<int, 1>{}) : go(I4{}, std::integral_constant<int, 1>{}); break;
            }
            return;
        }
        if (P >= kWideTileMinPixels) launch_conv<4, RELU, OUT_F32>(x, xs, xo, cin, w, bias, y, ys, yo, co, l);
        else launch_conv<2, RELU, OUT_F32>(x, xs, xo, cin, w, bias, y, ys, yo, co, l);
    }
    // LLGMI_UNET_TILE=0 runs every convolution with k_conv3x3 (the pixel-row kernel)
    static bool tile_conv_enabled() {
        const char *e = std::getenv("LLGMI_UNET_TILE");
        return !(e && e[0] == '0');
    }
    template <int TH, int NT, bool RELU, bool OUT_F32>
    void launch_conv_tile(const uint16_t *x, int xs, int xo, int cin, const uint16_t *w, const float *bias, void *y,
                          int ys, int yo, int co, int l) {
        constexpr size_t lds = conv_tile_lds<TH, NT>();
        auto kern = k_conv3x3_tile<TH, NT, RELU, OUT_F32>;
        {
            // the LDS attribute is set once per kernel and device
            static std::mutex mu;
            static std::set<int> done;
            int dev = 0;
            LLGMI_HIP(hipGetDevice(&dev));
            std::lock_guard<std::mutex> lock(mu);
            if (!done.count(dev)) {
                LLGMI_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                done.insert(dev);
            }
        }
        const int ntx = (wl(l) + kTileW - 1) / kTileW, nty = (hl(l) + TH - 1) / TH;
        const dim3 grid((unsigned)(ntx * nty * B), (unsigned)(co / (16 * NT)));
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, stream, x, xs, xo, cin, w, bias, y, ys, yo, co, B, hl(l), wl(l));
        LLGMI_LAUNCH_CHECK();
    }
    template <int MT, bool RELU, bool OUT_F32>
    void launch_conv(const uint16_t *x, int xs, int xo, int cin, const uint16_t *w, const float *bias, void *y,
                     int ys, int yo, int co, int l) {
        const int64_t P = pl(l);
        const int gx = (int)((P + 64 * MT - 1) / (64 * MT));
        auto lds = [&](int nt) { return (size_t)16 * nt * (cin + 8) * sizeof(uint16_t); };
        if (lds(co % 64 == 0 ? 4 : co % 32 == 0 ? 2 : 1) > kConvLdsMax)
            throw Error(LLGMI_ERR_INVALID, "internal: convolution weight slice exceeds LDS");
        if (co % 64 == 0)
            hipLaunchKernelGGL((k_conv3x3<MT, 4, RELU, OUT_F32>), dim3(gx, co / 64), dim3(256), lds(4), stream, x,
                               xs, xo, cin, w, bias, y, ys, yo, co, B, hl(l), wl(l));
        else if (co % 32 == 0)
            hipLaunchKernelGGL((k_conv3x3<MT, 2, RELU, OUT_F32>), dim3(gx, co / 32), dim3(256), lds(2), stream, x,
                               xs, xo, cin, w, bias, y, ys, yo, co, B, hl(l), wl(l));
        else
            hipLaunchKernelGGL((k_conv3x3<MT, 1, RELU, OUT_F32>), dim3(gx, co / 16), dim3(256), lds(1), stream, x,
                               xs, xo, cin, w, bias, y, ys, yo, co, B, hl(l), wl(l));
        LLGMI_LAUNCH_CHECK();
    }
    void wgrad(const uint16_t *dzp, int co, const uint16_t *x, int xs, int xo, int ci, const UNetConv &cv, int l) {
        if (co % 16 || ci % 16 || xo + ci > xs || co != cv.cout || ci != cv.cin)
            throw Error(LLGMI_ERR_INVALID, "internal: weight-gradient shape");
        const int64_t P = pl(l);
        float *dw = d_g + cv.w_off;
        const int mt = (co / 16) % 2 == 0 ? 2 : 1, nt = (ci / 16) % 2 == 0 ? 2 : 1;
        // pixel chunk per block: about 2048 blocks in all (each block ends in
        // 16*mt x 16*nt x 4 atomics), at least 2048 pixels, a multiple of 64
        const int tiles = (co / (16 * mt)) * (ci / (16 * nt));
        const int64_t nchunks = std::max<int64_t>(1, 2048 / (9 * tiles));
        const int chunk = (int)std::max<int64_t>(2048, ((P + nchunks - 1) / nchunks + 63) / 64 * 64);
        const int gx = (int)((P + chunk - 1) / chunk);
        const dim3 grid(gx, tiles, 9);
        if (mt == 2 && nt == 2)
            hipLaunchKernelGGL((k_conv3x3_wgrad<2, 2>), grid, dim3(256), 0, stream, dzp, co, x, xs, xo, ci, dw, B,
                               hl(l), wl(l), chunk);
        else if (mt == 2)
            hipLaunchKernelGGL((k_conv3x3_wgrad<2, 1>), grid, dim3(256), 0, stream, dzp, co, x, xs, xo, ci, dw, B,
                               hl(l), wl(l), chunk);
        else if (nt == 2)
            hipLaunchKernelGGL((k_conv3x3_wgrad<1, 2>), grid, dim3(256), 0, stream, dzp, co, x, xs, xo, ci, dw, B,
                               hl(l), wl(l), chunk);
        else
            hipLaunchKernelGGL((k_conv3x3_wgrad<1, 1>), grid, dim3(256), 0, stream, dzp, co, x, xs, xo, ci, dw, B,
                               hl(l), wl(l), chunk);
        const int bchunk = 1024;
        if (co > 256) throw Error(LLGMI_ERR_INVALID, "internal: bias gradient over 256 channels");
        if (co % 8 == 0)
            hipLaunchKernelGGL(k_bias_grad_v8, dim3((unsigned)((P + bchunk - 1) / bchunk)), dim3(256), 0, stream, dzp,
                               co, (int)P, bchunk, d_g + cv.b_off);
        else
            hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)((P + bchunk - 1) / bchunk)), dim3(256), 0, stream, dzp,
                               co, (int)P, bchunk, d_g + cv.b_off);
        LLGMI_LAUNCH_CHECK();
    }

    // conv index: enc l -> 2l, 2l+1; dec l (2..0) -> 8 + 2(2-l), +1; head -> 14
    const UNetConv &enc(int l, int k) const { return conv[(size_t)(2 * l + k)]; }
    const UNetConv &dec(int l, int k) const { return conv[(size_t)(2 * kLevels + 2 * (kLevels - 2 - l) + k)]; }
    const UNetConv &head() const { return conv.back(); }

    void forward(const uint16_t *x) {
        for (int l = 0; l < kLevels; ++l) {
            const uint16_t *src = l == 0 ? x : pool[l];
            const int cs = l == 0 ? 16 : c(l - 1);
            if (l > 0) {
                const int cprev = c(l - 1);
                const int ps = l - 1 < kLevels - 1 ? c(l) + cprev : cprev;   // eb[l-1] stride
                const int po = l - 1 < kLevels - 1 ? c(l) : 0;
                if (cprev % 8 == 0 && ps % 8 == 0 && po % 8 == 0)
                    hipLaunchKernelGGL(k_pool2_v8, dim3(grid_of(pl(l) * cprev / 8)), dim3(kNT), 0, stream, eb[l - 1],
                                       ps, po, cprev, pool[l], B, hl(l - 1), wl(l - 1));
                else
                    hipLaunchKernelGGL(k_pool2, dim3(grid_of(pl(l) * cprev)), dim3(kNT), 0, stream, eb[l - 1], ps, po,
                                       cprev, pool[l], B, hl(l - 1), wl(l - 1));
            }
            conv_fwd<true, false>(src, cs, 0, cs, enc(l, 0).wf, d_p + enc(l, 0).b_off, ea[l], c(l), 0, c(l), l);
            const int es = l < kLevels - 1 ? c(l + 1) + c(l) : c(l), eo = l < kLevels - 1 ? c(l + 1) : 0;
            conv_fwd<true, false>(ea[l], c(l), 0, c(l), enc(l, 1).wf, d_p + enc(l, 1).b_off, eb[l], es, eo, c(l), l);
        }
        for (int l = kLevels - 2; l >= 0; --l) {
            const uint16_t *below = l == kLevels - 2 ? eb[l + 1] : db[l + 1];
            const int cs = c(l + 1) + c(l);
            if (c(l + 1) % 8 == 0 && cs % 8 == 0)
                hipLaunchKernelGGL(k_up2_v8, dim3(grid_of(pl(l) * c(l + 1) / 8)), dim3(kNT), 0, stream, below,
                                   c(l + 1), cat[l], cs, 0, B, hl(l), wl(l));
            else
                hipLaunchKernelGGL(k_up2, dim3(grid_of(pl(l) * c(l + 1))), dim3(kNT), 0, stream, below, c(l + 1), cat[l],
                                   cs, 0, B, hl(l), wl(l));
            conv_fwd<true, false>(cat[l], cs, 0, cs, dec(l, 0).wf, d_p + dec(l, 0).b_off, da[l], c(l), 0, c(l), l);
            conv_fwd<true, false>(da[l], c(l), 0, c(l), dec(l, 1).wf, d_p + dec(l, 1).b_off, db[l], c(l), 0, c(l), l);
        }
        conv_fwd<false, true>(db[0], c(0), 0, c(0), head().wf, d_p + head().b_off, out, cout, 0, cout, 0);
        LLGMI_LAUNCH_CHECK();
    }

    // relu'(y) applied to g (dense C), cast to bf16 into dz[l]
    void relu_bwd(const float *g, const uint16_t *y, int ys, int yo, int C, int l) {
        if (C % 8 == 0 && (!y || (ys % 8 == 0 && yo % 8 == 0)))
            hipLaunchKernelGGL(k_relu_bwd_v8, dim3(grid_of(pl(l) * C / 8)), dim3(kNT), 0, stream, g, y, ys, yo, C, dz[l],
                               pl(l));
        else
            hipLaunchKernelGGL(k_relu_bwd, dim3(grid_of(pl(l) * C)), dim3(kNT), 0, stream, g, y, ys, yo, C, dz[l],
                               pl(l));
    }

    // gradient of the loss w.r.t. every parameter into d_g (overwritten)
    double loss_grad(const uint16_t *x, const uint16_t *tgt) {
        forward(x);
        LLGMI_HIP(hipMemsetAsync(d_g, 0, sizeof(float) * (size_t)nparam, stream));
        LLGMI_HIP(hipMemsetAsync(d_loss, 0, sizeof(double), stream));
        const double count = (double)B * std::min(wv, W) * std::min(hv, H) * cout_real;
        hipLaunchKernelGGL(k_mse, dim3(grid_of(pl(0) * cout)), dim3(kNT), 0, stream, out, tgt, cout, cout_real, wv,
                           hv, B, H, W, 1.0 / count, g_out, d_loss);
        LLGMI_HIP(hipMemcpyAsync(h_loss, d_loss, sizeof(double), hipMemcpyDeviceToHost, stream));
        // head
        relu_bwd(g_out, nullptr, 0, 0, cout, 0);
        wgrad(dz[0], cout, db[0], c(0), 0, c(0), head(), 0);
        conv_fwd<false, true>(dz[0], cout, 0, cout, head().wb, nullptr, g_b[0], c(0), 0, c(0), 0);
        // decoders, top level first
        for (int l = 0; l <= kLevels - 2; ++l) {
            const int cs = c(l + 1) + c(l);
            relu_bwd(g_b[l], db[l], c(l), 0, c(l), l);
            wgrad(dz[l], c(l), da[l], c(l), 0, c(l), dec(l, 1), l);
            conv_fwd<false, true>(dz[l], c(l), 0, c(l), dec(l, 1).wb, nullptr, g_a[l], c(l), 0, c(l), l);
            relu_bwd(g_a[l], da[l], c(l), 0, c(l), l);
            wgrad(dz[l], c(l), cat[l], cs, 0, cs, dec(l, 0), l);
            conv_fwd<false, true>(dz[l], c(l), 0, c(l), dec(l, 0).wb, nullptr, g_cat[l], cs, 0, cs, l);
            // up part -> gradient of the level below's decoder output (or the bottleneck)
            float *gbelow = l + 1 == kLevels - 1 ? g_eb[l + 1] : g_b[l + 1];
            hipLaunchKernelGGL(k_up2_bwd, dim3(grid_of(pl(l + 1) * c(l + 1))), dim3(kNT), 0, stream, g_cat[l], cs, 0,
                               c(l + 1), gbelow, B, hl(l), wl(l));
            // skip part -> gradient of enc_l's output (the pool gradient is added later)
            hipLaunchKernelGGL(k_slice_copy, dim3(grid_of(pl(l) * c(l))), dim3(kNT), 0, stream, g_cat[l], cs, c(l + 1),
                               c(l), g_eb[l], pl(l));
        }
        // encoders, bottom level first
        for (int l = kLevels - 1; l >= 0; --l) {
            const int es = l < kLevels - 1 ? c(l + 1) + c(l) : c(l), eo = l < kLevels - 1 ? c(l + 1) : 0;
            relu_bwd(g_eb[l], eb[l], es, eo, c(l), l);
            wgrad(dz[l], c(l), ea[l], c(l), 0, c(l), enc(l, 1), l);
            conv_fwd<false, true>(dz[l], c(l), 0, c(l), enc(l, 1).wb, nullptr, g_a[l], c(l), 0, c(l), l);
            relu_bwd(g_a[l], ea[l], c(l), 0, c(l), l);
            if (l == 0) {
                wgrad(dz[0], c(0), x, 16, 0, 16, enc(0, 0), 0);
            } else {
                wgrad(dz[l], c(l), pool[l], c(l - 1), 0, c(l - 1), enc(l, 0), l);
                conv_fwd<false, true>(dz[l], c(l), 0, c(l), enc(l, 0).wb, nullptr, g_pool[l], c(l - 1), 0, c(l - 1), l);
                hipLaunchKernelGGL(k_pool2_bwd, dim3(grid_of(pl(l - 1) * c(l - 1))), dim3(kNT), 0, stream, g_pool[l],
                                   c(l - 1), g_eb[l - 1], B, hl(l - 1), wl(l - 1));
            }
        }
        LLGMI_LAUNCH_CHECK();
        LLGMI_HIP(hipStreamSynchronize(stream));
        return *h_loss;
    }

    void adam(double lr) {
        ++t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)t), bc2 = 1.f - std::pow(b2, (float)t);
        hipLaunchKernelGGL(k_adam, dim3(grid_of(nparam)), dim3(kNT), 0, stream, d_p, d_g, d_m, d_v, nparam,
                           (float)lr, b1, b2, 1e-8f, bc1, bc2);
        cast_weights();
        LLGMI_LAUNCH_CHECK();
    }
};

}  // namespace llgmi

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using llgmi::Error;
using llgmi::UNet;

struct llgmi_ws {
    std::unique_ptr<UNet> net;
    int device = 0;
};

namespace {
int ws_fail(int code, const std::string &m) { return llgmi::set_last_error(code, m); }

void need(const void *p, int64_t bytes, const char *what) {
    const std::string e = llgmi::check_device_range(p, (size_t)bytes, what);
    if (!e.empty()) throw Error(LLGMI_ERR_INVALID, e);
}
int64_t x_bytes(const UNet &u) { return u.pl(0) * 16 * 2; }
int64_t t_bytes(const UNet &u) { return u.pl(0) * u.cout * 2; }
int64_t y_bytes(const UNet &u) { return u.pl(0) * u.cout * 4; }

template <typename F>
int ws_guard(llgmi_ws *ws, F &&f) {
    if (!ws || !ws->net) return ws_fail(LLGMI_ERR_INVALID, "warm-start network is NULL");
    try {
        LLGMI_HIP(hipSetDevice(ws->device));
        f(*ws->net);
        return LLGMI_OK;
    } catch (const Error &e) {
        return ws_fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return ws_fail(LLGMI_ERR_ALLOC, "host allocation failed");
    } catch (const std::exception &e) {
        return ws_fail(LLGMI_ERR_INVALID, e.what());
    }
}
}  // namespace

extern "C" {

int llgmi_ws_create(const llgmi_ws_shape *shape, uint64_t seed, int device, llgmi_ws **out) {
    if (!shape || !out) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    *out = nullptr;
    const llgmi_ws_shape &s = *shape;
    const int64_t div = (int64_t)1 << (llgmi::kLevels - 1);
    if (s.batch < 1 || s.height < div || s.width < div || s.height % div || s.width % div)
        return ws_fail(LLGMI_ERR_INVALID, "height and width must be positive multiples of 8, batch >= 1");
    if (s.in_channels < 1 || s.in_channels > 16) return ws_fail(LLGMI_ERR_INVALID, "in_channels must be 1..16");
    if (s.out_channels < 1 || s.out_channels > 256) return ws_fail(LLGMI_ERR_INVALID, "out_channels must be 1..256");
    if (s.base_channels < 16 || s.base_channels % 16 || s.base_channels > 32)
        return ws_fail(LLGMI_ERR_INVALID, "base_channels must be 16 or 32");
    if (s.batch * s.height * s.width > ((int64_t)1 << 28)) return ws_fail(LLGMI_ERR_INVALID, "batch too large");
    if (s.valid_width < 1 || s.valid_height < 1) return ws_fail(LLGMI_ERR_INVALID, "valid region is empty");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return ws_fail(LLGMI_ERR_INVALID, "device index out of range");
    std::unique_ptr<llgmi_ws> ws(new (std::nothrow) llgmi_ws());
    if (!ws) return ws_fail(LLGMI_ERR_ALLOC, "host allocation failed");
    ws->device = device;
    try {
        LLGMI_HIP(hipSetDevice(device));
        ws->net.reset(new UNet((int)s.batch, (int)s.height, (int)s.width, s.in_channels, s.out_channels,
                               s.base_channels, (int)s.valid_width, (int)s.valid_height, seed));
    } catch (const Error &e) {
        return ws_fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return ws_fail(LLGMI_ERR_ALLOC, "host allocation failed");
    }
    *out = ws.release();
    return LLGMI_OK;
}

int llgmi_ws_destroy(llgmi_ws *ws) {
    if (!ws) return LLGMI_OK;
    (void)hipSetDevice(ws->device);
    ws->net.reset();
    delete ws;
    return LLGMI_OK;
}

int llgmi_ws_param_count(llgmi_ws *ws, int64_t *n) {
    if (!n) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    return ws_guard(ws, [&](UNet &u) { *n = u.nparam; });
}

int llgmi_ws_get_params(llgmi_ws *ws, float *host) {
    if (!host) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    return ws_guard(ws, [&](UNet &u) {
        LLGMI_HIP(hipMemcpyAsync(host, u.d_p, sizeof(float) * (size_t)u.nparam, hipMemcpyDeviceToHost, u.stream));
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_set_params(llgmi_ws *ws, const float *host) {
    if (!host) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    return ws_guard(ws, [&](UNet &u) { u.set_params(host); });
}

int llgmi_ws_get_grads(llgmi_ws *ws, float *host) {
    if (!host) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    return ws_guard(ws, [&](UNet &u) {
        LLGMI_HIP(hipMemcpyAsync(host, u.d_g, sizeof(float) * (size_t)u.nparam, hipMemcpyDeviceToHost, u.stream));
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_forward(llgmi_ws *ws, const uint16_t *x_dev, float *y_dev) {
    if (!x_dev || !y_dev) return ws_fail(LLGMI_ERR_INVALID, "NULL buffer");
    return ws_guard(ws, [&](UNet &u) {
        need(x_dev, x_bytes(u), "x");
        need(y_dev, y_bytes(u), "y");
        u.forward(x_dev);
        const int64_t n = u.pl(0) * u.cout;
        LLGMI_HIP(hipMemcpyAsync(y_dev, u.out, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, u.stream));
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_loss_grad(llgmi_ws *ws, const uint16_t *x_dev, const uint16_t *t_dev, double *loss) {
    if (!x_dev || !t_dev || !loss) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    return ws_guard(ws, [&](UNet &u) {
        need(x_dev, x_bytes(u), "x");
        need(t_dev, t_bytes(u), "t");
        *loss = u.loss_grad(x_dev, t_dev);
    });
}

int llgmi_ws_adam_step(llgmi_ws *ws, double lr) {
    if (!(lr > 0.0)) return ws_fail(LLGMI_ERR_INVALID, "learning rate must be positive");
    return ws_guard(ws, [&](UNet &u) {
        u.adam(lr);
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_train_step(llgmi_ws *ws, const uint16_t *x_dev, const uint16_t *t_dev, double lr, double *loss) {
    if (!x_dev || !t_dev || !loss) return ws_fail(LLGMI_ERR_INVALID, "NULL argument");
    if (!(lr > 0.0)) return ws_fail(LLGMI_ERR_INVALID, "learning rate must be positive");
    return ws_guard(ws, [&](UNet &u) {
        need(x_dev, x_bytes(u), "x");
        need(t_dev, t_bytes(u), "t");
        *loss = u.loss_grad(x_dev, t_dev);
        u.adam(lr);
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_pack_input(llgmi_ws *ws, int64_t sample, const void *m_dev, int m_is_f32, int64_t nx, int64_t ny,
                        int64_t offx, int64_t offy, uint16_t *x_dev) {
    if (!m_dev || !x_dev) return ws_fail(LLGMI_ERR_INVALID, "NULL buffer");
    return ws_guard(ws, [&](UNet &u) {
        if (sample < 0 || sample >= u.B) throw Error(LLGMI_ERR_INVALID, "sample index out of range");
        if (nx < 1 || ny < 1 || offx < 0 || offy < 0 || offx + nx > u.W || offy + ny > u.H)
            throw Error(LLGMI_ERR_INVALID, "film does not fit the network's plane");
        need(m_dev, 3 * nx * ny * (m_is_f32 ? 4 : 8), "m");
        need(x_dev, x_bytes(u), "x");
        uint16_t *dst = x_dev + sample * (int64_t)u.H * u.W * 16;
        const int64_t n = (int64_t)u.H * u.W;
        if (m_is_f32)
            hipLaunchKernelGGL(llgmi::k_pack_input<float>, dim3(llgmi::grid_of(n)), dim3(256), 0, u.stream,
                               (const float *)m_dev, (int)nx, (int)ny, (int)offx, (int)offy, dst, u.H, u.W);
        else
            hipLaunchKernelGGL(llgmi::k_pack_input<double>, dim3(llgmi::grid_of(n)), dim3(256), 0, u.stream,
                               (const double *)m_dev, (int)nx, (int)ny, (int)offx, (int)offy, dst, u.H, u.W);
        LLGMI_LAUNCH_CHECK();
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_pack_phi(llgmi_ws *ws, int64_t sample, const double *phi_dev, int64_t px, int64_t py, int64_t pz,
                      double scale, uint16_t *t_dev) {
    if (!phi_dev || !t_dev) return ws_fail(LLGMI_ERR_INVALID, "NULL buffer");
    if (!(scale > 0.0)) return ws_fail(LLGMI_ERR_INVALID, "scale must be positive");
    return ws_guard(ws, [&](UNet &u) {
        if (sample < 0 || sample >= u.B) throw Error(LLGMI_ERR_INVALID, "sample index out of range");
        if (px < 1 || py < 1 || pz < 1 || px > u.W || py > u.H || pz > u.cout_real)
            throw Error(LLGMI_ERR_INVALID, "box does not fit the network's plane and channels");
        need(phi_dev, px * py * pz * 8, "phi");
        need(t_dev, t_bytes(u), "t");
        uint16_t *dst = t_dev + sample * (int64_t)u.H * u.W * u.cout;
        const int64_t n = (int64_t)u.H * u.W * u.cout;
        hipLaunchKernelGGL(llgmi::k_pack_phi, dim3(llgmi::grid_of(n)), dim3(256), 0, u.stream, phi_dev, (int)px,
                           (int)py, (int)pz, 1.0 / scale, dst, u.cout, u.H, u.W);
        LLGMI_LAUNCH_CHECK();
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

int llgmi_ws_unpack_phi(llgmi_ws *ws, int64_t sample, const float *y_dev, int64_t px, int64_t py, int64_t pz,
                        double scale, int accumulate, double *phi_dev) {
    if (!phi_dev || !y_dev) return ws_fail(LLGMI_ERR_INVALID, "NULL buffer");
    return ws_guard(ws, [&](UNet &u) {
        if (sample < 0 || sample >= u.B) throw Error(LLGMI_ERR_INVALID, "sample index out of range");
        if (px < 1 || py < 1 || pz < 1 || px > u.W || py > u.H || pz > u.cout_real)
            throw Error(LLGMI_ERR_INVALID, "box does not fit the network's plane and channels");
        need(phi_dev, px * py * pz * 8, "phi");
        need(y_dev, y_bytes(u), "y");
        const float *src = y_dev + sample * (int64_t)u.H * u.W * u.cout;
        const int64_t n = px * py * pz;
        hipLaunchKernelGGL(llgmi::k_unpack_phi, dim3(llgmi::grid_of(n)), dim3(256), 0, u.stream, src, u.cout, u.H,
                           u.W, scale, accumulate, phi_dev, (int)px, (int)py, (int)pz);
        LLGMI_LAUNCH_CHECK();
        LLGMI_HIP(hipStreamSynchronize(u.stream));
    });
}

}  // extern "C"

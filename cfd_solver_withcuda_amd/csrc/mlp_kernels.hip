// Kernels of the coordinate-MLP surrogate (warm start of the FP64 pseudo-time march): BF16
// weights and activations on gfx950 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation,
// fused bias/SiLU epilogues, split-K weight gradients with a fixed-order reduction, and Adam on
// FP32 master weights. The reference has no counterpart; the oracle's FP64 MLP (tests) is the
// checker.
//
// GEMM convention: C[r][c] = sum_k A[r][k] * B[c][k] with A [R x K] and B [C x K] both row-major
// (K contiguous), R, C multiples of 32 and K a multiple of 32. One wave computes a 32 x 32 tile
// of C as 2 x 2 MFMA 16x16 tiles (four such waves share their operands through LDS where R and C
// are multiples of 64). Operand lanes of v_mfma_f32_16x16x32_bf16: lane l holds
// A[l % 16][8 (l / 16) .. +8] and B[l % 16][8 (l / 16) .. +8]; result lane l holds
// C[4 (l / 16) + v][l % 16], v = 0..3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "mlp_kernels.h"

namespace cfd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    return *reinterpret_cast<const bf16x8 *>(p);
}

// Epilogues
enum { EPI_BIAS_SILU = MLP_EPI_BIAS_SILU, EPI_BIAS = MLP_EPI_BIAS, EPI_DSILU = MLP_EPI_DSILU, EPI_PARTIAL = MLP_EPI_PARTIAL };

// Epilogue of one wave's 32 x 32 tile at (r0, c0) of C (acc as the MFMA result lanes hold it);
// ks: the split-K range of a PARTIAL launch.
template <int EPI>
__device__ __forceinline__ void gemm_epilogue(const f32x4 (&acc)[2][2], int r0, int c0, int l, int ks, int R,
                                              int C, const float *__restrict__ bias,
                                              const void *__restrict__ Zprev, void *__restrict__ Zout,
                                              uint16_t *__restrict__ Y, uint16_t *__restrict__ Yt, int ldy,
                                              int ldyt, int c_valid) {
    const int lr = l % 16;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const int c = c0 + 16 * j + lr;
            const int rb = r0 + 16 * i + 4 * (l / 16);
            float v[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
            if constexpr (EPI == EPI_PARTIAL) {
                float *P = static_cast<float *>(Zout) + (size_t)ks * R * C;
                for (int q = 0; q < 4; q++) P[(size_t)(rb + q) * C + c] = v[q];
            } else if constexpr (EPI == EPI_BIAS) {
                float *Z = static_cast<float *>(Zout);
                const float bb = (c < c_valid) ? bias[c] : 0.f;
                for (int q = 0; q < 4; q++) Z[(size_t)(rb + q) * C + c] = (c < c_valid) ? v[q] + bb : 0.f;
            } else if constexpr (EPI == EPI_BIAS_SILU) {
                uint16_t *Zh = static_cast<uint16_t *>(Zout);
                const float bb = (c < c_valid) ? bias[c] : 0.f;
                uint16_t ys[4];
                for (int q = 0; q < 4; q++) {
                    const float z = (c < c_valid) ? v[q] + bb : 0.f;
                    Zh[(size_t)(rb + q) * C + c] = f2bf(z);
                    ys[q] = (c < c_valid) ? f2bf(z * sigmoidf_(z)) : (uint16_t)0;
                    Y[(size_t)(rb + q) * ldy + c] = ys[q];
                }
                // transposed copy: 4 consecutive rows of column c
                uint2 pk;
                pk.x = (uint32_t)ys[0] | ((uint32_t)ys[1] << 16);
                pk.y = (uint32_t)ys[2] | ((uint32_t)ys[3] << 16);
                *reinterpret_cast<uint2 *>(Yt + (size_t)c * ldyt + rb) = pk;
            } else {  // EPI_DSILU: dZprev = acc * silu'(Zprev), Zprev the bf16 pre-activations
                const uint16_t *Zh = static_cast<const uint16_t *>(Zprev);
                uint16_t ys[4];
                for (int q = 0; q < 4; q++) {
                    float d = 0.f;
                    if (c < c_valid) {
                        const float z = bf2f(Zh[(size_t)(rb + q) * C + c]);
                        const float s = sigmoidf_(z);
                        d = v[q] * (s * (1.0f + z * (1.0f - s)));
                    }
                    ys[q] = f2bf(d);
                    Y[(size_t)(rb + q) * ldy + c] = ys[q];
                }
                uint2 pk;
                pk.x = (uint32_t)ys[0] | ((uint32_t)ys[1] << 16);
                pk.y = (uint32_t)ys[2] | ((uint32_t)ys[3] << 16);
                *reinterpret_cast<uint2 *>(Yt + (size_t)c * ldyt + rb) = pk;
            }
        }
}

template <int EPI>
__global__ void __launch_bounds__(64) gemm_nt_kernel(const uint16_t *__restrict__ A, int lda,
                                                     const uint16_t *__restrict__ B, int ldb,
                                                     int R, int C, int K, int ksplit,
                                                     const float *__restrict__ bias,
                                                     const void *__restrict__ Zprev,
                                                     void *__restrict__ Zout,
                                                     uint16_t *__restrict__ Y, uint16_t *__restrict__ Yt,
                                                     int ldy, int ldyt, int c_valid) {
    const int tilesC = C / 32;
    const int tile = blockIdx.x;
    const int r0 = (tile / tilesC) * 32, c0 = (tile % tilesC) * 32;
    const int l = threadIdx.x;
    const int lr = l % 16, lk = 8 * (l / 16);
    // split-K: blockIdx.y selects the K range
    const int kper = K / ksplit;
    const int kb = blockIdx.y * kper, ke = kb + kper;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *a0 = A + (size_t)(r0 + lr) * lda + lk;
    const uint16_t *a1 = A + (size_t)(r0 + 16 + lr) * lda + lk;
    const uint16_t *b0 = B + (size_t)(c0 + lr) * ldb + lk;
    const uint16_t *b1 = B + (size_t)(c0 + 16 + lr) * ldb + lk;
    for (int k = kb; k < ke; k += 32) {
        const bf16x8 va0 = load8(a0 + k), va1 = load8(a1 + k);
        const bf16x8 vb0 = load8(b0 + k), vb1 = load8(b1 + k);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb1, acc[1][1], 0, 0, 0);
    }
    gemm_epilogue<EPI>(acc, r0, c0, l, blockIdx.y, R, C, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid);
}

// The same product for R and C multiples of 64: a workgroup of four waves computes a 64 x 64 block
// (wave w the 32 x 32 tile (w / 2, w % 2)), staging each 32-wide K slice of its 64 rows of A and B
// in LDS once (double-buffered, the next slice's global loads in flight during this slice's
// products) instead of every wave reading its operands from L2: half the L2 reads of the one-wave
// kernel. Same products in the same order per output, so the same results.
// Epilogue of a 64 x 64 block through LDS: the four waves' results (MFMA lane layout) go to a
// staged [64][65] float array S first, and each output is then written by whole 16-byte stores of
// consecutive elements: the fp32 rows of Z (or of the split-K partial), the bf16 rows of Y and the
// bf16 rows of Yt (columns of the block). Same values as gemm_epilogue.
constexpr int kStageRow = 65;

template <int EPI>
__device__ __forceinline__ void gemm64_epilogue(const f32x4 (&acc)[2][2], float *S, int R0, int C0, int w, int l,
                                                int t, int ks, int R, int C, const float *__restrict__ bias,
                                                const void *__restrict__ Zprev, void *__restrict__ Zout,
                                                uint16_t *__restrict__ Y, uint16_t *__restrict__ Yt, int ldy,
                                                int ldyt, int c_valid) {
    // 1. results into S (the loop's last barrier ended every read of the operand slices)
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const int c = 32 * (w % 2) + 16 * j + l % 16;
            const int r = 32 * (w / 2) + 16 * i + 4 * (l / 16);
            float bb = 0.f;
            bool cv = true;
            if constexpr (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS) {
                cv = C0 + c < c_valid;
                bb = bias[cv ? C0 + c : 0];
            }
            for (int q = 0; q < 4; q++) {
                float v = acc[i][j][q];
                if constexpr (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS) v = cv ? v + bb : 0.f;
                S[(r + q) * kStageRow + c] = v;
            }
        }
    __syncthreads();
    if constexpr (EPI == EPI_PARTIAL || EPI == EPI_BIAS || EPI == EPI_BIAS_SILU) {
        // 2. fp32 rows: four consecutive values per thread and pass
        for (int e = t; e < 64 * 16; e += 256) {
            const int r = e / 16, c = 4 * (e % 16);
            float *sr = S + r * kStageRow + c;
            const size_t o = (size_t)(R0 + r) * C + C0 + c;
            if constexpr (EPI == EPI_BIAS_SILU) {
                // bf16 pre-activations (what the SiLU' product reads back)
                const uint32_t lo = (uint32_t)f2bf(sr[0]) | ((uint32_t)f2bf(sr[1]) << 16);
                const uint32_t hi = (uint32_t)f2bf(sr[2]) | ((uint32_t)f2bf(sr[3]) << 16);
                *reinterpret_cast<uint2 *>(static_cast<uint16_t *>(Zout) + o) = make_uint2(lo, hi);
            } else {
                float *out = static_cast<float *>(Zout) + ((EPI == EPI_PARTIAL) ? (size_t)ks * R * C : 0);
                *reinterpret_cast<float4 *>(out + o) = make_float4(sr[0], sr[1], sr[2], sr[3]);
            }
            if constexpr (EPI == EPI_BIAS_SILU) {
                // the activation replaces the pre-activation in S (one SiLU per element)
                for (int q = 0; q < 4; q++) {
                    const float x = sr[q];
                    sr[q] = (C0 + c + q < c_valid) ? x * sigmoidf_(x) : 0.f;
                }
            }
        }
        if constexpr (EPI == EPI_BIAS_SILU) __syncthreads();
    }
    if constexpr (EPI == EPI_DSILU) {
        // 2. dZprev = acc * silu'(Zprev) in place in S, read by rows with the bf16 Zprev (padding
        // columns of Zprev hold zero; their results are selected to zero)
        for (int e = t; e < 64 * 16; e += 256) {
            const int r = e / 16, c = 4 * (e % 16);
            const uint2 z2 = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(Zprev) +
                                                              (size_t)(R0 + r) * C + C0 + c);
            const float zz[4] = {bf2f((uint16_t)(z2.x & 0xffff)), bf2f((uint16_t)(z2.x >> 16)),
                                 bf2f((uint16_t)(z2.y & 0xffff)), bf2f((uint16_t)(z2.y >> 16))};
            float *sr = S + r * kStageRow + c;
            for (int q = 0; q < 4; q++) {
                const float z = zz[q];
                const float sg = sigmoidf_(z);
                sr[q] = (C0 + c + q < c_valid) ? sr[q] * (sg * (1.0f + z * (1.0f - sg))) : 0.f;
            }
        }
        __syncthreads();
    }
    if constexpr (EPI == EPI_BIAS_SILU || EPI == EPI_DSILU) {
        // 3. bf16 rows of Y and of Yt: eight consecutive values per thread and pass
        auto val = [&](int r, int c) -> uint16_t { return f2bf(S[r * kStageRow + c]); };
        for (int e = t; e < 64 * 8; e += 256) {
            const int r = e / 8, c = 8 * (e % 8);
            uint32_t pk[4];
            for (int q = 0; q < 4; q++) pk[q] = (uint32_t)val(r, c + 2 * q) | ((uint32_t)val(r, c + 2 * q + 1) << 16);
            *reinterpret_cast<uint4 *>(Y + (size_t)(R0 + r) * ldy + C0 + c) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        }
        for (int e = t; e < 64 * 8; e += 256) {
            const int c = e / 8, r = 8 * (e % 8);
            uint32_t pk[4];
            for (int q = 0; q < 4; q++) pk[q] = (uint32_t)val(r + 2 * q, c) | ((uint32_t)val(r + 2 * q + 1, c) << 16);
            *reinterpret_cast<uint4 *>(Yt + (size_t)(C0 + c) * ldyt + R0 + r) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        }
    }
}

constexpr int kLdsRow = 40;  // bf16 per staged row: 32 + 8 of padding (conflict-free 16-byte reads)

template <int EPI>
__global__ void __launch_bounds__(256) gemm_nt64_kernel(const uint16_t *__restrict__ A, int lda,
                                                        const uint16_t *__restrict__ B, int ldb,
                                                        int R, int C, int K, int ksplit,
                                                        const float *__restrict__ bias,
                                                        const void *__restrict__ Zprev,
                                                        void *__restrict__ Zout,
                                                        uint16_t *__restrict__ Y, uint16_t *__restrict__ Yt,
                                                        int ldy, int ldyt, int c_valid) {
    // the K slices' operands; after the loop the same LDS stages the 64 x 64 result
    __shared__ __attribute__((aligned(16))) uint16_t sbuf[4 * 64 * kLdsRow];
    uint16_t(*sA)[64 * kLdsRow] = reinterpret_cast<uint16_t(*)[64 * kLdsRow]>(sbuf);
    uint16_t(*sB)[64 * kLdsRow] = reinterpret_cast<uint16_t(*)[64 * kLdsRow]>(sbuf + 2 * 64 * kLdsRow);
    const int tilesC = C / 64;
    const int R0 = (blockIdx.x / tilesC) * 64, C0 = (blockIdx.x % tilesC) * 64;
    const int t = threadIdx.x, w = t / 64, l = t % 64;
    const int lr = l % 16, lk = 8 * (l / 16);
    const int kper = K / ksplit;
    const int kb = blockIdx.y * kper, ke = kb + kper;
    // staging: thread t moves 8 bf16 of row t / 4 of each operand's 64 x 32 slice
    const int srow = t / 4, scol = 8 * (t % 4);
    const uint16_t *ga = A + (size_t)(R0 + srow) * lda + scol;
    const uint16_t *gb = B + (size_t)(C0 + srow) * ldb + scol;
    const int sofs = srow * kLdsRow + scol;
    *reinterpret_cast<bf16x8 *>(&sA[0][sofs]) = load8(ga + kb);
    *reinterpret_cast<bf16x8 *>(&sB[0][sofs]) = load8(gb + kb);
    __syncthreads();
    const int fa = (32 * (w / 2) + lr) * kLdsRow + lk;  // this lane's fragment rows in the slice
    const int fb = (32 * (w % 2) + lr) * kLdsRow + lk;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    int buf = 0;
    for (int k = kb; k < ke; k += 32) {
        const bool more = k + 32 < ke;
        bf16x8 na, nb;
        if (more) {
            na = load8(ga + k + 32);
            nb = load8(gb + k + 32);
        }
        const bf16x8 va0 = *reinterpret_cast<const bf16x8 *>(&sA[buf][fa]);
        const bf16x8 va1 = *reinterpret_cast<const bf16x8 *>(&sA[buf][fa + 16 * kLdsRow]);
        const bf16x8 vb0 = *reinterpret_cast<const bf16x8 *>(&sB[buf][fb]);
        const bf16x8 vb1 = *reinterpret_cast<const bf16x8 *>(&sB[buf][fb + 16 * kLdsRow]);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va0, vb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va1, vb1, acc[1][1], 0, 0, 0);
        if (more) {
            // the other buffer was last read before the previous barrier
            *reinterpret_cast<bf16x8 *>(&sA[buf ^ 1][sofs]) = na;
            *reinterpret_cast<bf16x8 *>(&sB[buf ^ 1][sofs]) = nb;
        }
        __syncthreads();
        buf ^= 1;
    }
    gemm64_epilogue<EPI>(acc, reinterpret_cast<float *>(sbuf), R0, C0, w, l, t, blockIdx.y, R, C, bias, Zprev, Zout,
                         Y, Yt, ldy, ldyt, c_valid);
}


// Normalised input features (n_in of them) of minibatch rows: row r takes sample idx[r] (or r
// when idx is null), padded with zeros to K0 columns; row-major X and transposed Xt.
__global__ void __launch_bounds__(256) encode_kernel(const float *__restrict__ feat, int n_in,
                                                     const int *__restrict__ idx, int rows,
                                                     const float *__restrict__ mean,
                                                     const float *__restrict__ scale, int K0,
                                                     uint16_t *__restrict__ X, uint16_t *__restrict__ Xt,
                                                     int ldxt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int s = idx ? idx[r] : r;
    // the row of X eight values (16 bytes) at a time; the columns of Xt one value per row, so
    // that neighbouring rows' threads write neighbouring bytes (K0 is a multiple of 32)
    for (int k0 = 0; k0 < K0; k0 += 8) {
        uint32_t pk[4];
        for (int q = 0; q < 8; q++) {
            const int k = k0 + q;
            uint16_t h = 0;
            if (k < n_in) h = f2bf((feat[(size_t)s * n_in + k] - mean[k]) * scale[k]);
            Xt[(size_t)k * ldxt + r] = h;
            pk[q / 2] = (q % 2) ? (pk[q / 2] | ((uint32_t)h << 16)) : (uint32_t)h;
        }
        *reinterpret_cast<uint4 *>(X + (size_t)r * K0 + k0) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    }
}

// Squared-error loss on normalised targets: dZ = 2 (Z - T) / (rows * n_out) for the first n_out
// columns (bf16, both layouts); per-row loss into lrow.
__global__ void __launch_bounds__(256) loss_kernel(const float *__restrict__ Z, int C,
                                                   const float *__restrict__ targ, int n_out,
                                                   const int *__restrict__ idx, int rows,
                                                   const float *__restrict__ mean,
                                                   const float *__restrict__ scale,
                                                   uint16_t *__restrict__ dZ, uint16_t *__restrict__ dZt,
                                                   int lddzt, float *__restrict__ lrow) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int s = idx ? idx[r] : r;
    const float inv = 2.0f / (float)(rows * n_out);
    float l = 0.f;
    // the row of dZ eight values (16 bytes) at a time (C is a multiple of 32), dZt as encode's Xt
    for (int c0 = 0; c0 < C; c0 += 8) {
        uint32_t pk[4];
        for (int q = 0; q < 8; q++) {
            const int c = c0 + q;
            float d = 0.f;
            if (c < n_out) {
                const float t = (targ[(size_t)s * n_out + c] - mean[c]) * scale[c];
                const float e = Z[(size_t)r * C + c] - t;
                l += e * e;
                d = inv * e;
            }
            const uint16_t h = f2bf(d);
            dZt[(size_t)c * lddzt + r] = h;
            pk[q / 2] = (q % 2) ? (pk[q / 2] | ((uint32_t)h << 16)) : (uint32_t)h;
        }
        *reinterpret_cast<uint4 *>(dZ + (size_t)r * C + c0) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    }
    lrow[r] = l / (float)n_out;
}




// Minibatch indices from a counter-based hash (SplitMix64) of (seed, step, row).
__global__ void __launch_bounds__(256) sample_kernel(unsigned long long seed, int step, int rows,
                                                     int n, int *__restrict__ idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)step * (unsigned long long)rows + r + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[r] = (int)(z % (unsigned long long)n);
}

// The minibatch of a captured training step's step *d_step (the step's Adam reads its rates
// from sched by the same index, and the step's last launch, the weight repacking, advances it).
__global__ void __launch_bounds__(256) sample_dev_kernel(unsigned long long seed, const int *__restrict__ d_step,
                                                         int rows, int n, int *__restrict__ idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int step = *d_step;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * ((unsigned long long)step * (unsigned long long)rows + r + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[r] = (int)(z % (unsigned long long)n);
}


// ---- all layers in one launch ---------------------------------------------------------------
// One Adam element update: moments m1, m2 and W -= lr (m1 c1) / (sqrt(m2 c2) + eps).
__device__ __forceinline__ void adam_one(float *W, float *m1, float *m2, const float *g, int i, float lr,
                                         float b1, float b2, float eps, float c1, float c2) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    W[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// Adam over weights [0, nw) and biases [nw, nw + nb) in one grid; with d_step the step's rates
// are sched[3 step .. 3 step + 2], else lr, c1, c2.
__global__ void __launch_bounds__(256) adam2_kernel(float *__restrict__ W, float *__restrict__ mW,
                                                    float *__restrict__ vW, const float *__restrict__ gW,
                                                    int nw, float *__restrict__ b, float *__restrict__ mb,
                                                    float *__restrict__ vb, const float *__restrict__ gb,
                                                    int nb, float lr, float c1, float c2,
                                                    const float *__restrict__ sched,
                                                    const int *__restrict__ d_step, float b1, float b2,
                                                    float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw + nb) return;
    if (d_step) {
        const int step = *d_step;
        lr = sched[3 * step];
        c1 = sched[3 * step + 1];
        c2 = sched[3 * step + 2];
    }
    if (i < nw) adam_one(W, mW, vW, gW, i, lr, b1, b2, eps, c1, c2);
    else adam_one(b, mb, vb, gb, i - nw, lr, b1, b2, eps, c1, c2);
}

// W -> Wb [Mp x Kp] and Wt [Kp x Mp] (bf16, zero padding) of layer blockIdx.y; block (0, 0) also
// advances the captured step's index when d_step is given (nothing in this launch reads it).
__global__ void __launch_bounds__(256) pack_layers_kernel(const float *__restrict__ W, MlpLayers t, int *d_step) {
    const int l = blockIdx.y;
    if (d_step && l == 0 && blockIdx.x == 0 && threadIdx.x == 0) *d_step += 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Mp = t.Mp[l], Kp = t.Kp[l], M = t.M[l], K = t.K[l];
    if (i >= Mp * Kp) return;
    const int m = i / Kp, k = i % Kp;
    const uint16_t h = (m < M && k < K) ? f2bf(W[t.woff[l] + (size_t)m * K + k]) : (uint16_t)0;
    t.Wb[l][i] = h;
    t.Wt[l][(size_t)k * Mp + m] = h;
}

// Split-K partial sums of layer blockIdx.y, in split order, into its unpadded gradient.
__global__ void __launch_bounds__(256) reduce_layers_kernel(const float *__restrict__ P, int S, MlpLayers t,
                                                            float *__restrict__ gW) {
    const int l = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int M = t.M[l], K = t.K[l], Kp = t.Kp[l];
    if (i >= M * K) return;
    const size_t MKp = (size_t)t.Mp[l] * Kp;
    const float *Pl = P + t.poff[l];
    const size_t o = (size_t)(i / K) * Kp + (i % K);
    float s = 0.f;
    for (int q = 0; q < S; q++) s += Pl[(size_t)q * MKp + o];
    gW[t.woff[l] + i] = s;
}

// Bias gradients of every layer: block b is row b of the layers' rows in turn; eight values per
// thread and step (one 16-byte load), four steps in flight (n a multiple of 512), the partial sums
// combined in a fixed order.
__global__ void __launch_bounds__(256) row_sum_layers_kernel(MlpLayers t, int n, float *__restrict__ gb) {
    __shared__ float red[256];
    int r = blockIdx.x, l = 0;
    while (l < t.L - 1 && r >= t.M[l]) r -= t.M[l++];
    const int tt = threadIdx.x;
    const uint16_t *row = t.dZt[l] + (size_t)r * n;
    float s4[4] = {0.f, 0.f, 0.f, 0.f};
    int j = 8 * tt;
    for (; j + 3 * 2048 < n; j += 4 * 2048)
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(row + j + u * 2048);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; q++) s4[u] += bf2f((uint16_t)(w[q] & 0xffff)) + bf2f((uint16_t)(w[q] >> 16));
        }
    for (; j < n; j += 2048) {
        const uint4 v = *reinterpret_cast<const uint4 *>(row + j);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; q++) s4[0] += bf2f((uint16_t)(w[q] & 0xffff)) + bf2f((uint16_t)(w[q] >> 16));
    }
    const float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    red[tt] = sum;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tt < h) red[tt] += red[tt + h];
        __syncthreads();
    }
    if (tt == 0) gb[t.boff[l] + r] = red[0];
}


}  // namespace

hipError_t mlp_sample_dev(unsigned long long seed, const int *d_step, int rows, int n, int *idx, hipStream_t s) {
    hipLaunchKernelGGL(sample_dev_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, seed, d_step, rows, n, idx);
    return hipGetLastError();
}

namespace {
int max_layer(const MlpLayers &t, bool padded) {
    int v = 0;
    for (int l = 0; l < t.L; l++) v = std::max(v, padded ? t.Mp[l] * t.Kp[l] : t.M[l] * t.K[l]);
    return v;
}
}  // namespace

hipError_t mlp_pack_layers(const float *W, const MlpLayers &t, int *d_step, hipStream_t s) {
    if (t.L < 1 || t.L > kMlpMaxLayers) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_layers_kernel, dim3((max_layer(t, true) + 255) / 256, t.L), dim3(256), 0, s, W, t, d_step);
    return hipGetLastError();
}

hipError_t mlp_reduce_layers(const float *P, int S, const MlpLayers &t, float *gW, hipStream_t s) {
    if (t.L < 1 || t.L > kMlpMaxLayers) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reduce_layers_kernel, dim3((max_layer(t, false) + 255) / 256, t.L), dim3(256), 0, s, P, S, t, gW);
    return hipGetLastError();
}

hipError_t mlp_row_sum_layers(const MlpLayers &t, int n, float *gb, hipStream_t s) {
    if (t.L < 1 || t.L > kMlpMaxLayers || n % 512) return hipErrorInvalidValue;
    int rows = 0;
    for (int l = 0; l < t.L; l++) rows += t.M[l];
    hipLaunchKernelGGL(row_sum_layers_kernel, dim3(rows), dim3(256), 0, s, t, n, gb);
    return hipGetLastError();
}

hipError_t mlp_adam2(float *W, float *mW, float *vW, const float *gW, int nw, float *b, float *mb, float *vb,
                     const float *gb, int nb, float lr, float b1, float b2, float eps, float c1, float c2,
                     hipStream_t s) {
    hipLaunchKernelGGL(adam2_kernel, dim3((nw + nb + 255) / 256), dim3(256), 0, s, W, mW, vW, gW, nw, b, mb, vb, gb, nb,
                       lr, c1, c2, nullptr, nullptr, b1, b2, eps);
    return hipGetLastError();
}

hipError_t mlp_adam2_dev(float *W, float *mW, float *vW, const float *gW, int nw, float *b, float *mb, float *vb,
                         const float *gb, int nb, const float *sched, const int *d_step, float b1, float b2,
                         float eps, hipStream_t s) {
    hipLaunchKernelGGL(adam2_kernel, dim3((nw + nb + 255) / 256), dim3(256), 0, s, W, mW, vW, gW, nw, b, mb, vb, gb, nb,
                       0.f, 0.f, 0.f, sched, d_step, b1, b2, eps);
    return hipGetLastError();
}

namespace {
template <int EPI>
void launch_gemm(const uint16_t *A, int lda, const uint16_t *B, int ldb, int R, int C, int K, int ksplit,
                 const float *bias, const void *Zprev, void *Zout, uint16_t *Y, uint16_t *Yt, int ldy, int ldyt,
                 int c_valid, hipStream_t s) {
    if (R % 64 == 0 && C % 64 == 0) {
        hipLaunchKernelGGL(gemm_nt64_kernel<EPI>, dim3((R / 64) * (C / 64), ksplit), dim3(256), 0, s, A, lda, B, ldb,
                           R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid);
    } else {
        hipLaunchKernelGGL(gemm_nt_kernel<EPI>, dim3((R / 32) * (C / 32), ksplit), dim3(64), 0, s, A, lda, B, ldb,
                           R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid);
    }
}
}  // namespace

hipError_t mlp_gemm(int epi, const uint16_t *A, int lda, const uint16_t *B, int ldb, int R, int C,
                    int K, int ksplit, const float *bias, const void *Zprev, void *Zout,
                    uint16_t *Y, uint16_t *Yt, int ldy, int ldyt, int c_valid, hipStream_t s) {
    if (R % 32 || C % 32 || K % 32 || ksplit < 1 || (K / 32) % ksplit) return hipErrorInvalidValue;
    switch (epi) {
    case EPI_BIAS_SILU:
        launch_gemm<EPI_BIAS_SILU>(A, lda, B, ldb, R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid, s);
        break;
    case EPI_BIAS:
        launch_gemm<EPI_BIAS>(A, lda, B, ldb, R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid, s);
        break;
    case EPI_DSILU:
        launch_gemm<EPI_DSILU>(A, lda, B, ldb, R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid, s);
        break;
    default:
        launch_gemm<EPI_PARTIAL>(A, lda, B, ldb, R, C, K, ksplit, bias, Zprev, Zout, Y, Yt, ldy, ldyt, c_valid, s);
        break;
    }
    return hipGetLastError();
}

hipError_t mlp_encode(const float *feat, int n_in, const int *idx, int rows, const float *mean,
                      const float *scale, int K0, uint16_t *X, uint16_t *Xt, int ldxt, hipStream_t s) {
    if (K0 % 8) return hipErrorInvalidValue;  // the kernel writes X rows 8 values at a time
    hipLaunchKernelGGL(encode_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, feat, n_in, idx, rows, mean, scale, K0, X, Xt, ldxt);
    return hipGetLastError();
}

hipError_t mlp_loss(const float *Z, int C, const float *targ, int n_out, const int *idx, int rows,
                    const float *mean, const float *scale, uint16_t *dZ, uint16_t *dZt, int lddzt,
                    float *lrow, hipStream_t s) {
    if (C % 8) return hipErrorInvalidValue;  // the kernel writes dZ rows 8 values at a time
    hipLaunchKernelGGL(loss_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, Z, C, targ, n_out, idx, rows, mean, scale, dZ, dZt, lddzt, lrow);
    return hipGetLastError();
}

hipError_t mlp_sample(unsigned long long seed, int step, int rows, int n, int *idx, hipStream_t s) {
    hipLaunchKernelGGL(sample_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, seed, step, rows, n, idx);
    return hipGetLastError();
}

}  // namespace cfd

// Launchers of the surrogate-MLP kernels (mlp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfd {

enum { MLP_EPI_BIAS_SILU = 0, MLP_EPI_BIAS = 1, MLP_EPI_DSILU = 2, MLP_EPI_PARTIAL = 3 };

// C[r][c] = sum_k A[r][k] B[c][k] (bf16 in, fp32 accumulate) with the epilogue `epi`:
//   BIAS_SILU: Zout = C + bias (fp32, ld C), Y = silu(Zout) (bf16, ld ldy), Yt = Y^T (ld ldyt)
//   BIAS:      Zout = C + bias
//   DSILU:     Y = C * silu'(Zprev) (bf16), Yt = Y^T
//   PARTIAL:   Zout[blockIdx.y] = C over the blockIdx.y-th of ksplit K ranges
// Columns c >= c_valid are written as zero. R, C, K multiples of 32; (K/32) % ksplit == 0.
hipError_t mlp_gemm(int epi, const uint16_t *A, int lda, const uint16_t *B, int ldb, int R, int C,
                    int K, int ksplit, const float *bias, const float *Zprev, float *Zout,
                    uint16_t *Y, uint16_t *Yt, int ldy, int ldyt, int c_valid, hipStream_t s);
hipError_t mlp_row_sum(const uint16_t *X, int ld, int rows, int n, float *out, hipStream_t s);
hipError_t mlp_encode(const float *feat, int n_in, const int *idx, int rows, const float *mean,
                      const float *scale, int K0, uint16_t *X, uint16_t *Xt, int ldxt, hipStream_t s);
hipError_t mlp_loss(const float *Z, int C, const float *targ, int n_out, const int *idx, int rows,
                    const float *mean, const float *scale, uint16_t *dZ, uint16_t *dZt, int lddzt,
                    float *lrow, hipStream_t s);
hipError_t mlp_adam(float *W, float *m1, float *m2, const float *g, int n, float lr, float b1,
                    float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t mlp_pack(const float *W, int M, int K, int Mp, int Kp, uint16_t *Wb, uint16_t *Wt, hipStream_t s);
// Sum of S partials [S][Mp x Kp] (MKp = Mp * Kp) over the valid [M x K] block, written unpadded.
hipError_t mlp_reduce_unpad(const float *P, int S, int MKp, int Kp, int M, int K, float *out, hipStream_t s);
hipError_t mlp_sample(unsigned long long seed, int step, int rows, int n, int *idx, hipStream_t s);
// Step-indexed forms for a captured training step: the step is *d_step on the device, Adam's
// learning rate and bias corrections are sched[3 step .. 3 step + 2], mlp_advance adds one to it.
hipError_t mlp_sample_dev(unsigned long long seed, const int *d_step, int rows, int n, int *idx, hipStream_t s);
hipError_t mlp_adam_dev(float *W, float *m1, float *m2, const float *g, int n, const float *sched,
                        const int *d_step, float b1, float b2, float eps, hipStream_t s);
hipError_t mlp_advance(int *d_step, hipStream_t s);

}  // namespace cfd

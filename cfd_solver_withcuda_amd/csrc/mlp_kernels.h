// Launchers of the surrogate-MLP kernels (mlp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfd {

enum { MLP_EPI_BIAS_SILU = 0, MLP_EPI_BIAS = 1, MLP_EPI_DSILU = 2, MLP_EPI_PARTIAL = 3 };

// Per-layer shapes and buffers of a network of up to kMlpMaxLayers weight layers, passed by value
// to the kernels that treat all layers in one launch.
constexpr int kMlpMaxLayers = 16;
struct MlpLayers {
    int L;
    int M[kMlpMaxLayers], K[kMlpMaxLayers], Mp[kMlpMaxLayers], Kp[kMlpMaxLayers];
    long long woff[kMlpMaxLayers], boff[kMlpMaxLayers], poff[kMlpMaxLayers];
    uint16_t *Wb[kMlpMaxLayers], *Wt[kMlpMaxLayers];
    const uint16_t *dZt[kMlpMaxLayers];
};

// C[r][c] = sum_k A[r][k] B[c][k] (bf16 in, fp32 accumulate) with the epilogue `epi`:
//   BIAS_SILU: Zout = C + bias (bf16 pre-activations, ld C), Y = silu(C + bias) (bf16, ld ldy),
//              Yt = Y^T (ld ldyt)
//   BIAS:      Zout = C + bias (fp32)
//   DSILU:     Y = C * silu'(Zprev) (Zprev the bf16 pre-activations; Y bf16), Yt = Y^T
//   PARTIAL:   Zout[blockIdx.y] = C over the blockIdx.y-th of ksplit K ranges (fp32)
// Columns c >= c_valid are written as zero. R, C, K multiples of 32; (K/32) % ksplit == 0.
hipError_t mlp_gemm(int epi, const uint16_t *A, int lda, const uint16_t *B, int ldb, int R, int C,
                    int K, int ksplit, const float *bias, const void *Zprev, void *Zout,
                    uint16_t *Y, uint16_t *Yt, int ldy, int ldyt, int c_valid, hipStream_t s);
hipError_t mlp_encode(const float *feat, int n_in, const int *idx, int rows, const float *mean,
                      const float *scale, int K0, uint16_t *X, uint16_t *Xt, int ldxt, hipStream_t s);
hipError_t mlp_loss(const float *Z, int C, const float *targ, int n_out, const int *idx, int rows,
                    const float *mean, const float *scale, uint16_t *dZ, uint16_t *dZt, int lddzt,
                    float *lrow, hipStream_t s);
hipError_t mlp_sample(unsigned long long seed, int step, int rows, int n, int *idx, hipStream_t s);
// Step-indexed form for a captured training step: the step is *d_step on the device.
hipError_t mlp_sample_dev(unsigned long long seed, const int *d_step, int rows, int n, int *idx, hipStream_t s);
// All layers in one launch each (t.L <= kMlpMaxLayers):
//   W -> Wb [Mp x Kp] and Wt [Kp x Mp] (bf16, zero padding) for every layer; with d_step,
//   *d_step += 1 as well (the end of a captured step);
hipError_t mlp_pack_layers(const float *W, const MlpLayers &t, int *d_step, hipStream_t s);
//   gW[woff + m K + k] = sum over the S split-K partials [S][Mp x Kp] of layer l at P + poff[l];
hipError_t mlp_reduce_layers(const float *P, int S, const MlpLayers &t, float *gW, hipStream_t s);
//   gb[boff + m] = sum of the first n entries of row m of dZt[l] (leading dimension n), m < M[l];
hipError_t mlp_row_sum_layers(const MlpLayers &t, int n, float *gb, hipStream_t s);
//   Adam on the weights [nw] and the biases [nb] together: with lr, c1, c2 (mlp_adam2), or with
//   the learning rate and bias corrections sched[3 step .. 3 step + 2] of step *d_step.
hipError_t mlp_adam2(float *W, float *mW, float *vW, const float *gW, int nw, float *b, float *mb, float *vb,
                     const float *gb, int nb, float lr, float b1, float b2, float eps, float c1, float c2,
                     hipStream_t s);
hipError_t mlp_adam2_dev(float *W, float *mW, float *vW, const float *gW, int nw, float *b, float *mb, float *vb,
                         const float *gb, int nb, const float *sched, const int *d_step, float b1, float b2,
                         float eps, hipStream_t s);

}  // namespace cfd

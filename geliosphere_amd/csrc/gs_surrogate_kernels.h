// gs_surrogate_kernels.h - launchers of the surrogate's gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// GEMM epilogues
enum {
    EPI_STORE = 0,   // Cf = acc
    EPI_HIDDEN = 1,  // Zf = acc + bias, Cb = bf16(silu(Zf))
    EPI_OUT = 2,     // Cf = acc + bias
    EPI_DSILU = 3    // Cf = acc * silu'(Zf), Cb = bf16(Cf)
};

struct GemmArgs {
    const __bf16 *A;
    const __bf16 *B;
    int M, N, K;
    int lda, ldb, ldc;
    size_t strideA, strideB, strideC;   // per ensemble member (elements)
    const float *bias;
    size_t strideBias;
    float *Zf;                          // EPI_HIDDEN: written; EPI_DSILU: read
    float *Cf;
    __bf16 *Cb;
};

// Fused inference (all layers of a member per 128-row tile, activations in LDS):
// hidden width 256, input and output padded to 32, 2..kMlpMaxLayers weight layers.
constexpr int kMlpMaxLayers = 8;
struct MlpArgs {
    const __bf16 *X;                    // [rows][32] standardised inputs
    int rows, L;                        // L weight layers
    const __bf16 *Pb;                   // bf16 parameters, member stride Pm
    const float *P;                     // fp32 parameters (biases)
    size_t Pm;
    size_t woff[kMlpMaxLayers], boff[kMlpMaxLayers];
    float *Out;                         // [members][rows][32]
};

hipError_t sg_launch_gemm(bool a_col, bool b_col, int epi, const GemmArgs &g, int members, hipStream_t s);
hipError_t sg_launch_mlp_forward(const MlpArgs &g, int members, hipStream_t s);
hipError_t sg_launch_colsum(const float *X, int rows, int cols, size_t strideX, float *out, size_t strideOut,
                            int members, hipStream_t s);
hipError_t sg_launch_loss(const float *Y, const float *T, const float *W, int n, size_t strideY, size_t strideT,
                          float *wsum, float *dYf, __bf16 *dYb, float *loss_part, int parts, int members,
                          hipStream_t s);
hipError_t sg_launch_gather(const __bf16 *X, const float *Y, const float *W, const uint32_t *idx, int rows, int in_w,
                            int out_w, __bf16 *bx, float *by, float *bw, int members, hipStream_t s);
hipError_t sg_launch_adam(float *P, const float *G, float *Mo, float *Ve, __bf16 *Pb, size_t n, float lr, float b1,
                          float b2, float eps, float wd, float c1, float c2, hipStream_t s);
hipError_t sg_launch_to_bf16(const float *x, __bf16 *y, size_t n, hipStream_t s);
// Standardised, zero-padded bf16 rows [n][ldo] from fp32 rows [n][ni]:
// y = bf16((x - mu) / sd), with ni <= 32.
hipError_t sg_launch_standardise(const float *x, int n, int ni, int ldo, const float *mu, const float *sd, __bf16 *y,
                                 hipStream_t s);
// Fallback flags per query from the ensemble spread sd [n][no] and the raw
// inputs x [n][ni]: 0x1 spread above limit on an output of mask, 0x2 input
// outside [lo, hi]; no, ni <= 32.
hipError_t sg_launch_flags(const float *sd, const float *x, int n, int no, int ni, uint32_t mask, float limit,
                           const float *lo, const float *hi, uint8_t *flags, hipStream_t s);
hipError_t sg_launch_stats(const float *Y, int E, int n, int ldy, int cols, size_t strideY, float *mean, float *sd,
                           hipStream_t s);

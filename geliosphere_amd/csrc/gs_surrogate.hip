// gs_surrogate.hip - MLP-ensemble surrogate kernels for gfx950 (BF16 MFMA,
// FP32 accumulate): forward, backward, loss, Adam, ensemble statistics.
//
// One GEMM kernel serves every product of training and inference:
//   C[m][n] = sum_k A(m,k) B(n,k)
// with A(m,k) stored row-major (A[m*lda + k]) or column-major (A[k*lda + m])
// and likewise B(n,k), so that
//   forward        Z = X W^T        A = X (row),  B = W (row)
//   data gradient  dX = dZ W        A = dZ (row), B(n=i,k=o) = W[o][i] (col)
//   weight grad    dW = dZ^T X      A(m=o,k=b) = dZ[b][o] (col), B(n=i,k=b) = X[b][i] (col)
// Tiles: a 256-thread workgroup owns a 64x64 block of C; each wave a 32x32
// quarter as 2x2 v_mfma_f32_16x16x32_bf16 tiles.  K advances 32 at a time
// through LDS (rows padded to 40 bf16 so the 16-byte fragment reads of a
// wave spread over the banks).  Fragment layout of the 16x16x32 MFMA
// (checked on the device): lane l holds A[l%16][8(l/16)..8(l/16)+7],
// B[8(l/16)..+7][l%16], and result D[4(l/16)+v][l%16], v = 0..3.
// The ensemble index is blockIdx.z; every operand has a per-member stride.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_surrogate_kernels.h"

namespace gs {
namespace sg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32, LDS_ROW = TK + 8;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <bool A_COL, bool B_COL, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __bf16 sA[TM * LDS_ROW];
    __shared__ __bf16 sB[TN * LDS_ROW];
    const int e = blockIdx.z;
    const __bf16 *A = g.A + (size_t)e * g.strideA;
    const __bf16 *B = g.B + (size_t)e * g.strideB;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += TK) {
        // ---- global -> LDS, 8 bf16 per thread per operand, zero outside M/N
        if (!A_COL) {
            const int r = t >> 2, c = (t & 3) * 8, m = m0 + r;
            bf16x8 v = {};
            if (m < g.M) v = *(const bf16x8 *)(A + (size_t)m * g.lda + k0 + c);
            *(bf16x8 *)(sA + r * LDS_ROW + c) = v;
        } else {
            const int kk = t >> 3, c = (t & 7) * 8;
            const __bf16 *src = A + (size_t)(k0 + kk) * g.lda + m0 + c;
#pragma unroll
            for (int i = 0; i < 8; ++i) sA[(c + i) * LDS_ROW + kk] = (m0 + c + i < g.M) ? src[i] : (__bf16)0.0f;
        }
        if (!B_COL) {
            const int r = t >> 2, c = (t & 3) * 8, n = n0 + r;
            bf16x8 v = {};
            if (n < g.N) v = *(const bf16x8 *)(B + (size_t)n * g.ldb + k0 + c);
            *(bf16x8 *)(sB + r * LDS_ROW + c) = v;
        } else {
            const int kk = t >> 3, c = (t & 7) * 8;
            const __bf16 *src = B + (size_t)(k0 + kk) * g.ldb + n0 + c;
#pragma unroll
            for (int i = 0; i < 8; ++i) sB[(c + i) * LDS_ROW + kk] = (n0 + c + i < g.N) ? src[i] : (__bf16)0.0f;
        }
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8 *)(sA + (wm + 16 * i + fr) * LDS_ROW + fk);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8 *)(sB + (wn + 16 * j + fr) * LDS_ROW + fk);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // ---- epilogue
    const size_t eo = (size_t)e * g.strideC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const size_t o = eo + (size_t)m * g.ldc + n;
                const float x = acc[i][j][v];
                if (EPI == EPI_HIDDEN) {
                    const float z = x + g.bias[(size_t)e * g.strideBias + n];
                    if (g.Zf) g.Zf[o] = z;   // kept for the backward pass only
                    g.Cb[o] = (__bf16)silu(z);
                } else if (EPI == EPI_OUT) {
                    g.Cf[o] = x + g.bias[(size_t)e * g.strideBias + n];
                } else if (EPI == EPI_DSILU) {
                    const float d = x * dsilu(g.Zf[o]);
                    g.Cf[o] = d;
                    g.Cb[o] = (__bf16)d;
                } else {   // EPI_STORE
                    g.Cf[o] = x;
                }
            }
}

// Fused inference of one member over a tile of FR rows: every layer runs in
// one workgroup, with the bf16 activations held in LDS between layers instead
// of making a round trip through HBM.  Hidden width 256, padded input and
// output widths 32.  Each output element sees the same MFMA sequence as in
// gemm_kernel (one 16x16x32 step per 32 values of k, k ascending) and the
// same epilogue, so the outputs equal the layer-by-layer forward pass.
constexpr int FR = 128, FH = 256, FLD = FH + 8, FRT = FR / 16;

// One hidden layer in place: every wave reads the whole tile, then (after a
// barrier) overwrites its own 64 columns with the activations.
template <int K>
__device__ __forceinline__ void mlp_hidden_layer(__bf16 *buf, const __bf16 *W, const float *bias, int lane,
                                                 int wave) {
    const int fr = lane & 15, fk = (lane >> 4) * 8, n0 = wave * 64;
    f32x4 acc[FRT][4];
#pragma unroll
    for (int i = 0; i < FRT; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 b[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *(const bf16x8 *)(W + (size_t)(n0 + 16 * j + fr) * K + k0 + fk);
#pragma unroll
        for (int i = 0; i < FRT; ++i) {
            const bf16x8 a = *(const bf16x8 *)(buf + (16 * i + fr) * FLD + k0 + fk);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = n0 + 16 * j + fr;
        const float bn = bias[n];
#pragma unroll
        for (int i = 0; i < FRT; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = 16 * i + 4 * (lane >> 4) + v;
                buf[m * FLD + n] = (__bf16)silu(acc[i][j][v] + bn);
            }
    }
}

__global__ void __launch_bounds__(256, 2) mlp_forward_kernel(MlpArgs g) {
    __shared__ __bf16 act[FR * FLD];
    const int e = blockIdx.y, row0 = blockIdx.x * FR;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const __bf16 *Pb = g.Pb + (size_t)e * g.Pm;
    const float *P = g.P + (size_t)e * g.Pm;
    for (int q = t; q < FR * 4; q += 256) {   // input tile: FR rows x 32 bf16, zero beyond the last row
        const int r = q >> 2, c = (q & 3) * 8;
        bf16x8 v = {};
        if (row0 + r < g.rows) v = *(const bf16x8 *)(g.X + (size_t)(row0 + r) * 32 + c);
        *(bf16x8 *)(act + r * FLD + c) = v;
    }
    __syncthreads();
    mlp_hidden_layer<32>(act, Pb + g.woff[0], P + g.boff[0], lane, wave);
    __syncthreads();
    for (int l = 1; l < g.L - 1; ++l) {
        mlp_hidden_layer<FH>(act, Pb + g.woff[l], P + g.boff[l], lane, wave);
        __syncthreads();
    }
    // output layer, 32 padded columns: wave w owns rows FR/4 * w .. FR/4 * (w + 1) - 1
    constexpr int ORT = FRT / 4;
    const int l = g.L - 1, fr = lane & 15, fk = (lane >> 4) * 8, r0 = wave * (FR / 4);
    const __bf16 *W = Pb + g.woff[l];
    const float *bias = P + g.boff[l];
    f32x4 acc[ORT][2];
#pragma unroll
    for (int i = 0; i < ORT; ++i) acc[i][0] = acc[i][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < FH; k0 += 32) {
        bf16x8 b[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8 *)(W + (size_t)(16 * j + fr) * FH + k0 + fk);
#pragma unroll
        for (int i = 0; i < ORT; ++i) {
            const bf16x8 a = *(const bf16x8 *)(act + (r0 + 16 * i + fr) * FLD + k0 + fk);
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < ORT; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = r0 + 16 * i + 4 * (lane >> 4) + v, n = 16 * j + fr;
                if (row0 + m < g.rows) g.Out[((size_t)e * g.rows + row0 + m) * 32 + n] = acc[i][j][v] + bias[n];
            }
}

// Column sums of a [rows x cols] fp32 matrix per member (bias gradients).
// One thread per column, rows walked in order: deterministic.
__global__ void colsum_kernel(const float *X, int rows, int cols, size_t strideX, float *out, size_t strideOut) {
    const int e = blockIdx.y;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const float *x = X + (size_t)e * strideX;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += x[(size_t)r * cols + c];
    out[(size_t)e * strideOut + c] = s;
}

// Weighted squared error: L = sum w (y - t)^2 / sum w per member.
// dY = 2 w (y - t) / sum w, in fp32 and bf16.  wsum[e] holds sum w.
__global__ void loss_grad_kernel(const float *Y, const float *T, const float *W, int n, size_t strideY,
                                 size_t strideT, const float *wsum, float *dYf, __bf16 *dYb, float *loss_part) {
    const int e = blockIdx.y;
    __shared__ float red[256];
    float l = 0.f;
    const float inv = 1.0f / fmaxf(wsum[e], 1e-30f);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float y = Y[(size_t)e * strideY + i], t = T[(size_t)e * strideT + i], w = W[(size_t)e * strideT + i];
        const float d = y - t;
        l += w * d * d * inv;
        const float gr = 2.0f * w * d * inv;
        dYf[(size_t)e * strideY + i] = gr;
        dYb[(size_t)e * strideY + i] = (__bf16)gr;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[(size_t)e * gridDim.x + blockIdx.x] = red[0];
}

// Sum of the loss weights of each member's batch.
__global__ void wsum_kernel(const float *W, int n, size_t strideT, float *wsum) {
    const int e = blockIdx.x;
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += W[(size_t)e * strideT + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) wsum[e] = red[0];
}

// Gather a member's minibatch: rows idx[e][r] of the resident data set.
__global__ void gather_kernel(const __bf16 *X, const float *Y, const float *W, const uint32_t *idx, int rows,
                              int in_w, int out_w, __bf16 *bx, float *by, float *bw) {
    const int e = blockIdx.y;
    const int r = blockIdx.x;
    if (r >= rows) return;
    const uint32_t s = idx[(size_t)e * rows + r];
    for (int c = threadIdx.x; c < in_w; c += blockDim.x)
        bx[((size_t)e * rows + r) * in_w + c] = X[(size_t)s * in_w + c];
    for (int c = threadIdx.x; c < out_w; c += blockDim.x) {
        by[((size_t)e * rows + r) * out_w + c] = Y[(size_t)s * out_w + c];
        bw[((size_t)e * rows + r) * out_w + c] = W[(size_t)s * out_w + c];
    }
}

// Adam (decoupled weight decay) over the flat parameter vector of every
// member; refreshes the bf16 copy the GEMMs read.
__global__ void adam_kernel(float *P, const float *G, float *Mo, float *Ve, __bf16 *Pb, size_t n, float lr,
                            float b1, float b2, float eps, float wd, float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float g = G[i];
        const float m = b1 * Mo[i] + (1.0f - b1) * g;
        const float v = b2 * Ve[i] + (1.0f - b2) * g * g;
        Mo[i] = m;
        Ve[i] = v;
        float p = P[i];
        p -= lr * ((m / c1) / (sqrtf(v / c2) + eps) + wd * p);
        P[i] = p;
        Pb[i] = (__bf16)p;
    }
}

__global__ void to_bf16_kernel(const float *x, __bf16 *y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = (__bf16)x[i];
}

struct StdArgs {
    float mu[32], sd[32];
};

// One thread per padded element.  The subtraction and the IEEE-rounded
// division are those the host would do, so the rows equal a host
// standardisation followed by the bf16 conversion.
__global__ void standardise_kernel(const float *x, int n, int ni, int ldo, StdArgs a, __bf16 *y) {
    const size_t total = (size_t)n * ldo;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / ldo;
        const int c = (int)(i - r * ldo);
        float v = 0.0f;
        if (c < ni) v = (x[r * ni + c] - a.mu[c]) / a.sd[c];
        y[i] = (__bf16)v;
    }
}

struct FlagArgs {
    float lo[32], hi[32];
};

// Fallback flags, one thread per query: bit 0x1 when an output selected by
// `mask` has a spread that is not <= `limit` (NaN included), bit 0x2 when an
// input lies outside [lo, hi] (NaN included).
__global__ void fallback_flags_kernel(const float *sd, const float *x, int n, int no, int ni, uint32_t mask,
                                      float limit, FlagArgs b, uint8_t *flags) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t f = 0;
    for (int c = 0; c < no; ++c)
        if (((mask >> c) & 1u) && !(sd[(size_t)r * no + c] <= limit)) f |= 0x1u;
    for (int c = 0; c < ni; ++c) {
        const float v = x[(size_t)r * ni + c];
        if (!(v >= b.lo[c] && v <= b.hi[c])) f |= 0x2u;
    }
    flags[r] = (uint8_t)f;
}

// Mean and (population) standard deviation over the members of [E][n][ldy]
// outputs, first `cols` columns.
__global__ void ensemble_stats_kernel(const float *Y, int E, int n, int ldy, int cols, size_t strideY, float *mean,
                                      float *sd) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n * cols) return;
    const size_t r = i / cols, c = i % cols;
    float s = 0.f, s2 = 0.f;
    for (int e = 0; e < E; ++e) {
        const float y = Y[(size_t)e * strideY + r * ldy + c];
        s += y;
        s2 += y * y;
    }
    const float mu = s / E;
    mean[i] = mu;
    sd[i] = sqrtf(fmaxf(s2 / E - mu * mu, 0.f));
}

}  // namespace sg
}  // namespace gs

using namespace gs::sg;

template <bool AC, bool BC, int EPI>
static hipError_t launch_gemm_t(const GemmArgs &g, int members, hipStream_t s) {
    const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, members);
    hipLaunchKernelGGL((gemm_kernel<AC, BC, EPI>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t sg_launch_gemm(bool a_col, bool b_col, int epi, const GemmArgs &g, int members, hipStream_t s) {
    if (g.K % TK != 0 || g.M <= 0 || g.N <= 0 || members <= 0) return hipErrorInvalidValue;
    // the column-major loaders read 8 consecutive m (n) values per thread
    if ((a_col && (g.lda % 8)) || (b_col && (g.ldb % 8)) || (!a_col && (g.lda % 8)) || (!b_col && (g.ldb % 8)))
        return hipErrorInvalidValue;
    if (!a_col && !b_col) {
        if (epi == EPI_HIDDEN) return launch_gemm_t<false, false, EPI_HIDDEN>(g, members, s);
        if (epi == EPI_OUT) return launch_gemm_t<false, false, EPI_OUT>(g, members, s);
        return launch_gemm_t<false, false, EPI_STORE>(g, members, s);
    }
    if (!a_col && b_col) {
        if (epi == EPI_DSILU) return launch_gemm_t<false, true, EPI_DSILU>(g, members, s);
        return launch_gemm_t<false, true, EPI_STORE>(g, members, s);
    }
    if (a_col && b_col) return launch_gemm_t<true, true, EPI_STORE>(g, members, s);
    return hipErrorInvalidValue;
}

hipError_t sg_launch_mlp_forward(const MlpArgs &g, int members, hipStream_t s) {
    if (g.rows <= 0 || members <= 0 || g.L < 2 || g.L > kMlpMaxLayers) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((g.rows + FR - 1) / FR), members), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t sg_launch_colsum(const float *X, int rows, int cols, size_t strideX, float *out, size_t strideOut,
                            int members, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 255) / 256, members), dim3(256), 0, s, X, rows, cols, strideX,
                       out, strideOut);
    return hipGetLastError();
}

hipError_t sg_launch_loss(const float *Y, const float *T, const float *W, int n, size_t strideY, size_t strideT,
                          float *wsum, float *dYf, __bf16 *dYb, float *loss_part, int parts, int members,
                          hipStream_t s) {
    hipLaunchKernelGGL(wsum_kernel, dim3(members), dim3(256), 0, s, W, n, strideT, wsum);
    hipLaunchKernelGGL(loss_grad_kernel, dim3(parts, members), dim3(256), 0, s, Y, T, W, n, strideY, strideT, wsum,
                       dYf, dYb, loss_part);
    return hipGetLastError();
}

hipError_t sg_launch_gather(const __bf16 *X, const float *Y, const float *W, const uint32_t *idx, int rows, int in_w,
                            int out_w, __bf16 *bx, float *by, float *bw, int members, hipStream_t s) {
    hipLaunchKernelGGL(gather_kernel, dim3(rows, members), dim3(64), 0, s, X, Y, W, idx, rows, in_w, out_w, bx, by,
                       bw);
    return hipGetLastError();
}

hipError_t sg_launch_adam(float *P, const float *G, float *Mo, float *Ve, __bf16 *Pb, size_t n, float lr, float b1,
                          float b2, float eps, float wd, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3(1024), dim3(256), 0, s, P, G, Mo, Ve, Pb, n, lr, b1, b2, eps, wd, c1, c2);
    return hipGetLastError();
}

hipError_t sg_launch_to_bf16(const float *x, __bf16 *y, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t sg_launch_standardise(const float *x, int n, int ni, int ldo, const float *mu, const float *sd, __bf16 *y,
                                 hipStream_t s) {
    if (ni < 1 || ni > 32 || ni > ldo) return hipErrorInvalidValue;
    StdArgs a{};
    for (int c = 0; c < ni; ++c) {
        a.mu[c] = mu[c];
        a.sd[c] = sd[c];
    }
    hipLaunchKernelGGL(standardise_kernel, dim3(1024), dim3(256), 0, s, x, n, ni, ldo, a, y);
    return hipGetLastError();
}

hipError_t sg_launch_flags(const float *sd, const float *x, int n, int no, int ni, uint32_t mask, float limit,
                           const float *lo, const float *hi, uint8_t *flags, hipStream_t s) {
    if (n <= 0 || no < 0 || no > 32 || ni < 1 || ni > 32) return hipErrorInvalidValue;
    FlagArgs b{};
    for (int c = 0; c < ni; ++c) {
        b.lo[c] = lo[c];
        b.hi[c] = hi[c];
    }
    hipLaunchKernelGGL(fallback_flags_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sd, x, n, no, ni,
                       mask, limit, b, flags);
    return hipGetLastError();
}

hipError_t sg_launch_stats(const float *Y, int E, int n, int ldy, int cols, size_t strideY, float *mean, float *sd,
                           hipStream_t s) {
    const size_t total = (size_t)n * cols;
    hipLaunchKernelGGL(ensemble_stats_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Y, E, n, ldy,
                       cols, strideY, mean, sd);
    return hipGetLastError();
}

// closure_kernels.hpp -- device kernels of the learned spectral-convolution
// closure: one generic batched, strided GEMM on the CDNA4 matrix cores (BF16
// operands, FP32 accumulation, optional split-K), and the pointwise pieces
// (bias, GELU, loss, reductions, Adam).  The network itself is composed of
// these on the host (closure_api.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace shrt {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// C[z](i, j) = alpha * sum_k A[z](i, k) * B[z](k, j) + beta * C[z](i, j), every
// operand addressed through its own strides (so transposes and interleaved
// views cost nothing).  With ksplit > 1 the K range is cut into ksplit parts
// whose partial products go to `work` ([z][split][M][N], row-major) and are
// combined by k_gemm_reduce.
struct GemmArgs {
    const float* A;
    const float* B;
    float* C;
    float* work;
    int M, N, K, batch, ksplit, kchunk;
    long long sAz, sAi, sAk, sBz, sBk, sBj, sCz, sCi, sCj;
    float alpha, beta;
};

constexpr int kBM = 64, kBN = 64, kBK = 32, kLdsK = kBK + 8;   // padded rows (80 bytes) against bank conflicts

// 256 threads: four wavefronts, each a 32x32 quarter of the 64x64 tile as 2x2
// MFMA 16x16x32 blocks (gfx950's BF16 matrix instruction with K = 32).  Per K
// step every thread stages eight consecutive k of one A row and of one B
// column as BF16 in LDS ([m][k] and [n][k]); a lane's MFMA operand is then one
// 16-byte LDS read: A[row = lane%16][k = 8*(lane/16)..+7],
// B[k = 8*(lane/16)..+7][col = lane%16]; the result D[4*(lane/16) + r][lane%16]
// sits in register r.
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[kBM * kLdsK];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[kBN * kLdsK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int n0 = blockIdx.x * kBN, m0 = blockIdx.y * kBM;
    const int z = blockIdx.z / g.ksplit, split = blockIdx.z - z * g.ksplit;
    const int kbeg = split * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
    const float* A = g.A + z * g.sAz;
    const float* B = g.B + z * g.sBz;
    const int lr = t >> 2, lk = (t & 3) * 8;    // staging: row/col lr, k lk..lk+7
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        u16x8 av, bv;
        const int am = m0 + lr, bn = n0 + lr;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + lk + u;
            av[u] = to_bf16((am < g.M && k < kend) ? A[am * g.sAi + k * g.sAk] : 0.f);
            bv[u] = to_bf16((bn < g.N && k < kend) ? B[k * g.sBk + bn * g.sBj] : 0.f);
        }
        __syncthreads();   // the previous step's operands have been read
        *reinterpret_cast<u16x8*>(As + lr * kLdsK + lk) = av;
        *reinterpret_cast<u16x8*>(Bs + lr * kLdsK + lk) = bv;
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(As + (wm + i * 16 + fr) * kLdsK + fk));
            b[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(Bs + (wn + i * 16 + fr) * kLdsK + fk));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }

    const int col = lane & 15, rbase = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + rbase + r, cc = n0 + wn + j * 16 + col;
                if (row >= g.M || cc >= g.N) continue;
                if (g.ksplit > 1) {
                    g.work[(((long long)z * g.ksplit + split) * g.M + row) * g.N + cc] = acc[i][j][r];
                } else {
                    float* c = g.C + z * g.sCz + row * g.sCi + cc * g.sCj;
                    *c = g.beta != 0.f ? g.alpha * acc[i][j][r] + g.beta * *c : g.alpha * acc[i][j][r];
                }
            }
}

// C = alpha * sum over splits of work + beta * C (fixed summation order).
__global__ void k_gemm_reduce(GemmArgs g) {
    const long long mn = (long long)g.M * g.N, total = mn * g.batch;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(e / mn);
        const long long r = e - z * mn;
        const int row = (int)(r / g.N), cc = (int)(r - (long long)row * g.N);
        const float* p = g.work + (long long)z * g.ksplit * mn + r;
        float s = 0.f;
        for (int q = 0; q < g.ksplit; ++q) s += p[q * mn];
        float* c = g.C + z * g.sCz + row * g.sCi + cc * g.sCj;
        *c = g.beta != 0.f ? g.alpha * s + g.beta * *c : g.alpha * s;
    }
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
__device__ __forceinline__ float gelu(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    const float du = 0.7978845608028654f * (1.f + 3.f * 0.044715f * x * x);
    return 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * du;
}

// z[b][c][p] += bias[c]; optionally h = gelu(z).  n = points per channel.
__global__ void k_bias_act(float* z, float* h, const float* bias, int C, long long n, long long total, int act) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)((e / n) % C);
        const float v = z[e] + bias[c];
        z[e] = v;
        if (act) h[e] = gelu(v);
    }
}

// dz = dh * gelu'(z)
// The (0, 0) mode's cosine coefficient of `rows` coefficient rows scaled by s
// (the separable transforms carry sqrt(2)/N for every mode; that mode's basis
// column is 1/N).
__global__ void k_scale_first(float* base, long long pitch, int rows, float s) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) base[(long long)r * pitch] *= s;
}

__global__ void k_gelu_back(const float* dh, const float* z, float* dz, long long total) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x)
        dz[e] = dh[e] * gelu_grad(z[e]);
}

// Wavefront sum of a double.
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Block-wide sum of a double (blockDim.x = 256); result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    __syncthreads();
    return s;
}

// out[c] (+)= sum over b, p of x[b][c][p]; one workgroup per channel, FP64 sums.
__global__ void __launch_bounds__(256) k_channel_sum(const float* x, float* out, int B, int C, long long n, int accumulate) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = 0; b < B; ++b) {
        const float* p = x + ((long long)b * C + c) * n;
        for (long long i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[c] = accumulate ? out[c] + (float)s : (float)s;
}

// out[i] (+)= sum over b of parts[b][i]
__global__ void k_sum_parts(const float* parts, float* out, int B, long long n, int accumulate) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < B; ++b) s += parts[(long long)b * n + i];
        out[i] = accumulate ? out[i] + (float)s : (float)s;
    }
}

// Mean squared error over n values: partial sums per workgroup in FP64 and the
// gradient d = 2 (out - tgt) / n.
__global__ void __launch_bounds__(256) k_mse(const float* out, const float* tgt, float* d, double* partial, long long n) {
    __shared__ double red[4];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float e = out[i] - tgt[i];
        s += (double)e * e;
        d[i] = 2.f * e / (float)n;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int n, double scale, double* out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) *out = s * scale;
}

// Adam (bias-corrected step size lr_t computed on the host).
__global__ void k_adam(float* p, const float* g, float* m, float* v, long long n, float lr_t, float b1, float b2,
                       float eps) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr_t * mi / (sqrtf(vi) + eps);
    }
}

// Gather samples idx[0..B) of a dataset [S][len] into a batch [B][len].
__global__ void k_gather(const float* data, const int* idx, float* out, int B, long long len) {
    const long long total = (long long)B * len;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e / len);
        out[e] = data[(long long)idx[b] * len + (e - (long long)b * len)];
    }
}

}  // namespace nn
}  // namespace shrt

// closure_api.hip -- the learned spectral-convolution closure (no reference
// counterpart): network state, forward and backward passes composed of the
// MFMA GEMM and pointwise kernels of closure_kernels.hpp, Adam training on a
// device-resident data set, and the C ABI of include/shrt.h.
//
// Network on an N x N grid (physical layout f[ix*N + iy]):
//   h0 = P x + p                                (lift: 4 inputs -> C channels)
//   h_{l+1} = gelu(W_l h_l + w_l + K_l h_l),  l < L
//   out = Q h_L + q                             (projection: C -> 2 outputs)
// K_l is the spectral convolution on the retained modes |kx| <= kmax,
// 0 <= ky <= kmax (half plane): X = h E (analysis on an orthonormal cos/sin
// basis E), per mode Yc = Xc a^T - Xs b^T, Ys = Xc b^T + Xs a^T (a complex
// weight a + ib per mode and channel pair, so K_l commutes with translations),
// K_l h = Y E^T.  Every contraction is a BF16 MFMA GEMM with FP32 accumulation.
//
// E is never formed: its columns s*cos(a + b), s*sin(a + b), a = 2 pi kx ix/N,
// b = 2 pi ky iy/N, factor into a y stage (the cos b, sin b of ky = 0..kmax)
// and an x stage (s*cos a, s*sin a of kx = -kmax..kmax), so the analysis is
// two truncated-DFT GEMMs of O(N^2 kmax) work per channel instead of one of
// O(N^2 kmax^2), and the synthesis is their transpose.  Coefficient rows are
// stored as [ky][kx = -kmax..kmax][cos, sin] (pitch DP = (kmax+1)(4 kmax+2)):
// the modes in the order above start at offset OFF = 2 kmax, the ky = 0,
// kx < 0 entries before it are not modes (zero where a synthesis reads them).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shrt.h"
#include "closure_kernels.hpp"

using namespace shrt::nn;

namespace shrt {
void set_last_error(const std::string& msg);   // shrt_api.hip: the text of shrt_last_error
}

namespace {

shrt_status cfail(shrt_status s, const std::string& m) {
    shrt::set_last_error(m);
    return s;
}

#define NN_TRY(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) return cfail(SHRT_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct shrt_closure {
    int N = 0, kmax = 0, C = 0, L = 0, Mm = 0, D = 0, maxB = 0;
    long long NP = 0, np = 0;
    // parameter offsets (floats) in the flat buffer
    long long oP = 0, op = 0, oQ = 0, oq = 0;
    std::vector<long long> oa, ob, oW, ow;
    hipStream_t stream = nullptr;
    int K1 = 0, J = 0, DP = 0, OFF = 0;   // ky count, kx-row length (cos, sin), coefficient pitch, mode offset
    float* Fy = nullptr;    // [N][2*K1]: cos b, sin b
    float* Fxc = nullptr;   // [N][J]: s cos a, s sin a
    float* Fxs = nullptr;   // [N][J]: -s sin a, s cos a
    float* T = nullptr;     // [maxB*C*N][2*K1]: the y stage of an analysis or the x stage of a synthesis
    float *prm = nullptr, *grd = nullptr, *am = nullptr, *av = nullptr;
    float *x = nullptr, *out = nullptr, *dout = nullptr, *tgt = nullptr;
    std::vector<float*> h, z, X;
    float *Y = nullptr, *dY = nullptr, *dX = nullptr, *dh = nullptr, *dz = nullptr, *parts = nullptr;
    float* work = nullptr;
    long long work_len = 0;
    double* partial = nullptr;
    double* loss = nullptr;
    int npartial = 1024;
    // training data
    float *din = nullptr, *dtg = nullptr;
    int nsamples = 0;
    int* idx = nullptr;
    long long adam_t = 0;
    std::vector<void*> allocs;
    // ordering against a solver stream (closure_forward_on)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

shrt_status nn_alloc(shrt_closure* n, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return cfail(SHRT_ERR_ALLOC, "closure: device allocation failed");
    n->allocs.push_back(*p);
    return SHRT_OK;
}

dim3 pw_grid(long long total) {
    long long b = (total + 255) / 256;
    return dim3((unsigned)(b < 2048 ? (b < 1 ? 1 : b) : 2048));
}

// One batched strided GEMM (see GemmArgs); picks a K split when the output
// tiles alone cannot fill the chip.
hipError_t gemm(shrt_closure* n, const float* A, long long sAz, long long sAi, long long sAk, const float* B,
                long long sBz, long long sBk, long long sBj, float* C, long long sCz, long long sCi, long long sCj,
                int M, int N, int K, int batch, float alpha, float beta) {
    GemmArgs g{};
    g.A = A; g.B = B; g.C = C; g.work = n->work;
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.sAz = sAz; g.sAi = sAi; g.sAk = sAk; g.sBz = sBz; g.sBk = sBk; g.sBj = sBj;
    g.sCz = sCz; g.sCi = sCi; g.sCj = sCj;
    g.alpha = alpha; g.beta = beta;
    const long long tiles = (long long)((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN) * batch;
    int ks = 1;
    if (tiles < 512 && K > 512) {
        ks = (int)std::min<long long>(K / 256, std::max<long long>(1, 1024 / tiles));
        ks = std::max(1, std::min(ks, 256));
        while (ks > 1 && (long long)batch * ks * M * N > n->work_len) ks /= 2;
    }
    int chunk = (K + ks - 1) / ks;
    chunk = (chunk + kBK - 1) / kBK * kBK;
    ks = (K + chunk - 1) / chunk;
    g.ksplit = ks;
    g.kchunk = chunk;
    const dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, batch * ks);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, n->stream, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ks == 1) return e;
    hipLaunchKernelGGL(k_gemm_reduce, pw_grid((long long)M * N * batch), dim3(256), 0, n->stream, g);
    return hipGetLastError();
}

// The retained modes in a fixed order: ky = 0 with kx = 0..kmax, then for each
// ky = 1..kmax the kx = -kmax..kmax.
void mode_list(int kmax, std::vector<int>& kx, std::vector<int>& ky) {
    kx.clear();
    ky.clear();
    for (int y = 0; y <= kmax; ++y)
        for (int x = (y == 0 ? 0 : -kmax); x <= kmax; ++x) {
            kx.push_back(x);
            ky.push_back(y);
        }
}

// Analysis of R rows h[r][ix*N + iy] (row pitch NP) into coefficient rows
// X[r] (pitch DP, modes from OFF): X = h E.
hipError_t analysis(shrt_closure* n, const float* h, float* X, int R) {
    const int N = n->N, K2 = 2 * n->K1;
    // y stage: T[(r, ix)][(ky, c)] = sum_iy h[r][ix][iy] (cos b, sin b)
    hipError_t e = gemm(n, h, 0, N, 1, n->Fy, 0, K2, 1, n->T, 0, K2, 1, R * N, K2, N, 1, 1.f, 0.f);
    if (e != hipSuccess) return e;
    // x stage per row r: X[r][ky][kx, cs] = sum_ix Tc Fxc + Ts Fxs
    const long long tz = (long long)N * K2;
    if ((e = gemm(n, n->T, tz, 2, K2, n->Fxc, 0, n->J, 1, X, n->DP, n->J, 1, n->K1, n->J, N, R, 1.f, 0.f))) return e;
    if ((e = gemm(n, n->T + 1, tz, 2, K2, n->Fxs, 0, n->J, 1, X, n->DP, n->J, 1, n->K1, n->J, N, R, 1.f, 1.f))) return e;
    hipLaunchKernelGGL(k_scale_first, dim3((R + 255) / 256), dim3(256), 0, n->stream, X + n->OFF, (long long)n->DP, R,
                       (float)M_SQRT1_2);
    return hipGetLastError();
}

// Synthesis of R coefficient rows Y[r] (pitch DP; the entries before OFF
// zero) into z[r][ix*N + iy] = (Y E^T)[r] + beta z.  Scales Y's (0, 0)
// cosine in place.
hipError_t synthesis(shrt_closure* n, float* Y, float* z, int R, float beta) {
    const int N = n->N, K2 = 2 * n->K1;
    hipLaunchKernelGGL(k_scale_first, dim3((R + 255) / 256), dim3(256), 0, n->stream, Y + n->OFF, (long long)n->DP, R,
                       (float)M_SQRT1_2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // x stage: T[(r, ix)][(ky, c)] = sum_{kx, cs} Y[r][ky][kx, cs] (Fxc, Fxs)[ix][kx, cs]
    const long long tz = (long long)N * K2;
    if ((e = gemm(n, Y, n->DP, n->J, 1, n->Fxc, 0, 1, n->J, n->T, tz, 2, K2, n->K1, N, n->J, R, 1.f, 0.f))) return e;
    if ((e = gemm(n, Y, n->DP, n->J, 1, n->Fxs, 0, 1, n->J, n->T + 1, tz, 2, K2, n->K1, N, n->J, R, 1.f, 0.f))) return e;
    // y stage: z[(r, ix)][iy] = sum_{ky, c} T (cos b, sin b)
    return gemm(n, n->T, 0, K2, 1, n->Fy, 0, 1, K2, z, 0, N, 1, R * N, N, K2, 1, 1.f, beta);
}

hipError_t forward(shrt_closure* n, int B) {
    const int C = n->C, D = n->DP, Mm = n->Mm;   // D: the coefficient row pitch
    const long long NP = n->NP;
    const float* P = n->prm + n->oP;
    hipError_t e = gemm(n, P, 0, 4, 1, n->x, 4 * NP, NP, 1, n->h[0], C * NP, NP, 1, C, (int)NP, 4, B, 1.f, 0.f);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bias_act, pw_grid(B * C * NP), dim3(256), 0, n->stream, n->h[0], nullptr, n->prm + n->op, C,
                       NP, B * C * NP, 0);
    for (int l = 0; l < n->L; ++l) {
        const float* a = n->prm + n->oa[l];
        const float* b = n->prm + n->ob[l];
        // analysis X = h E
        if ((e = analysis(n, n->h[l], n->X[l], B * C))) return e;
        // per mode: Yc = Xc a^T - Xs b^T, Ys = Xc b^T + Xs a^T
        const float* Xc = n->X[l] + n->OFF;
        const float* Xs = n->X[l] + n->OFF + 1;
        float* Yc = n->Y + n->OFF;
        const long long xi = (long long)C * D;
        if ((e = gemm(n, Xc, 2, xi, D, a, (long long)C * C, 1, C, Yc, 2, xi, D, B, C, C, Mm, 1.f, 0.f))) return e;
        if ((e = gemm(n, Xs, 2, xi, D, b, (long long)C * C, 1, C, Yc, 2, xi, D, B, C, C, Mm, -1.f, 1.f))) return e;
        if ((e = gemm(n, Xc, 2, xi, D, b, (long long)C * C, 1, C, Yc + 1, 2, xi, D, B, C, C, Mm, 1.f, 0.f))) return e;
        if ((e = gemm(n, Xs, 2, xi, D, a, (long long)C * C, 1, C, Yc + 1, 2, xi, D, B, C, C, Mm, 1.f, 1.f))) return e;
        // synthesis z = Y E^T, then z += W h
        if ((e = synthesis(n, n->Y, n->z[l], B * C, 0.f))) return e;
        if ((e = gemm(n, n->prm + n->oW[l], 0, C, 1, n->h[l], C * NP, NP, 1, n->z[l], C * NP, NP, 1, C, (int)NP, C, B,
                      1.f, 1.f)))
            return e;
        hipLaunchKernelGGL(k_bias_act, pw_grid(B * C * NP), dim3(256), 0, n->stream, n->z[l], n->h[l + 1],
                           n->prm + n->ow[l], C, NP, B * C * NP, 1);
    }
    if ((e = gemm(n, n->prm + n->oQ, 0, C, 1, n->h[n->L], C * NP, NP, 1, n->out, 2 * NP, NP, 1, 2, (int)NP, C, B, 1.f,
                  0.f)))
        return e;
    hipLaunchKernelGGL(k_bias_act, pw_grid(B * 2 * NP), dim3(256), 0, n->stream, n->out, nullptr, n->prm + n->oq, 2,
                       NP, B * 2 * NP, 0);
    return hipGetLastError();
}

// Gradient of the loss whose output gradient is in dout; parameters' gradients to grd.
hipError_t backward(shrt_closure* n, int B) {
    const int C = n->C, D = n->DP, Mm = n->Mm;   // D: the coefficient row pitch
    const long long NP = n->NP;
    float* G = n->grd;
    hipError_t e;
    // projection
    if ((e = gemm(n, n->dout, 2 * NP, NP, 1, n->h[n->L], C * NP, 1, NP, n->parts, 2 * C, C, 1, 2, C, (int)NP, B, 1.f,
                  0.f)))
        return e;
    hipLaunchKernelGGL(k_sum_parts, pw_grid(2 * C), dim3(256), 0, n->stream, n->parts, G + n->oQ, B, (long long)2 * C, 0);
    hipLaunchKernelGGL(k_channel_sum, dim3(2), dim3(256), 0, n->stream, n->dout, G + n->oq, B, 2, NP, 0);
    if ((e = gemm(n, n->prm + n->oQ, 0, 1, C, n->dout, 2 * NP, NP, 1, n->dh, C * NP, NP, 1, C, (int)NP, 2, B, 1.f, 0.f)))
        return e;
    for (int l = n->L - 1; l >= 0; --l) {
        const float* a = n->prm + n->oa[l];
        const float* b = n->prm + n->ob[l];
        hipLaunchKernelGGL(k_gelu_back, pw_grid(B * C * NP), dim3(256), 0, n->stream, n->dh, n->z[l], n->dz,
                           B * C * NP);
        hipLaunchKernelGGL(k_channel_sum, dim3(C), dim3(256), 0, n->stream, n->dz, G + n->ow[l], B, C, NP, 0);
        // pointwise weights and the pointwise part of dh
        if ((e = gemm(n, n->dz, C * NP, NP, 1, n->h[l], C * NP, 1, NP, n->parts, (long long)C * C, C, 1, C, C, (int)NP,
                      B, 1.f, 0.f)))
            return e;
        hipLaunchKernelGGL(k_sum_parts, pw_grid((long long)C * C), dim3(256), 0, n->stream, n->parts, G + n->oW[l], B,
                           (long long)C * C, 0);
        if ((e = gemm(n, n->prm + n->oW[l], 0, 1, C, n->dz, C * NP, NP, 1, n->dh, C * NP, NP, 1, C, (int)NP, C, B, 1.f,
                      0.f)))
            return e;
        // spectral part: dY = dz E
        if ((e = analysis(n, n->dz, n->dY, B * C))) return e;
        const long long xi = (long long)C * D, cc = (long long)C * C;
        const float* dYc = n->dY + n->OFF;
        const float* dYs = n->dY + n->OFF + 1;
        float* dXc = n->dX + n->OFF;
        // dXc = dYc a + dYs b, dXs = -dYc b + dYs a
        if ((e = gemm(n, dYc, 2, xi, D, a, cc, C, 1, dXc, 2, xi, D, B, C, C, Mm, 1.f, 0.f))) return e;
        if ((e = gemm(n, dYs, 2, xi, D, b, cc, C, 1, dXc, 2, xi, D, B, C, C, Mm, 1.f, 1.f))) return e;
        if ((e = gemm(n, dYc, 2, xi, D, b, cc, C, 1, dXc + 1, 2, xi, D, B, C, C, Mm, -1.f, 0.f))) return e;
        if ((e = gemm(n, dYs, 2, xi, D, a, cc, C, 1, dXc + 1, 2, xi, D, B, C, C, Mm, 1.f, 1.f))) return e;
        // da = dYc^T Xc + dYs^T Xs, db = -dYc^T Xs + dYs^T Xc (sums over the batch)
        const float* Xc = n->X[l] + n->OFF;
        const float* Xs = n->X[l] + n->OFF + 1;
        float* ga = G + n->oa[l];
        float* gb = G + n->ob[l];
        if ((e = gemm(n, dYc, 2, D, xi, Xc, 2, xi, D, ga, cc, C, 1, C, C, B, Mm, 1.f, 0.f))) return e;
        if ((e = gemm(n, dYs, 2, D, xi, Xs, 2, xi, D, ga, cc, C, 1, C, C, B, Mm, 1.f, 1.f))) return e;
        if ((e = gemm(n, dYc, 2, D, xi, Xs, 2, xi, D, gb, cc, C, 1, C, C, B, Mm, -1.f, 0.f))) return e;
        if ((e = gemm(n, dYs, 2, D, xi, Xc, 2, xi, D, gb, cc, C, 1, C, C, B, Mm, 1.f, 1.f))) return e;
        // dh += dX E^T
        if ((e = synthesis(n, n->dX, n->dh, B * C, 1.f))) return e;
    }
    // lift
    if ((e = gemm(n, n->dh, C * NP, NP, 1, n->x, 4 * NP, 1, NP, n->parts, (long long)C * 4, 4, 1, C, 4, (int)NP, B, 1.f,
                  0.f)))
        return e;
    hipLaunchKernelGGL(k_sum_parts, pw_grid((long long)C * 4), dim3(256), 0, n->stream, n->parts, G + n->oP, B,
                       (long long)C * 4, 0);
    hipLaunchKernelGGL(k_channel_sum, dim3(C), dim3(256), 0, n->stream, n->dh, G + n->op, B, C, NP, 0);
    return hipGetLastError();
}

hipError_t loss_and_grad(shrt_closure* n, int B) {
    const long long cnt = (long long)B * 2 * n->NP;
    hipLaunchKernelGGL(k_mse, dim3(n->npartial), dim3(256), 0, n->stream, n->out, n->tgt, n->dout, n->partial, cnt);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, n->stream, n->partial, n->npartial, 1.0 / (double)cnt,
                       n->loss);
    return hipGetLastError();
}

}  // namespace

namespace shrt {

// Device-side access for the in-place correction of a solver context
// (shrt_closure_correct): grid size, the input [4][N*N] and output [2][N*N]
// buffers of sample 0, and one forward pass ordered after the work enqueued on
// the solver stream `s`, with `s` ordered after it.
int closure_grid(const shrt_closure* n) { return n->N; }
float* closure_input(shrt_closure* n) { return n->x; }
const float* closure_output(const shrt_closure* n) { return n->out; }

hipError_t closure_forward_on(shrt_closure* n, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (!n->ev_in) e = hipEventCreateWithFlags(&n->ev_in, hipEventDisableTiming);
    if (e == hipSuccess && !n->ev_out) e = hipEventCreateWithFlags(&n->ev_out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(n->ev_in, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(n->stream, n->ev_in, 0);
    if (e == hipSuccess) e = forward(n, 1);
    if (e == hipSuccess) e = hipEventRecord(n->ev_out, n->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, n->ev_out, 0);
    return e;
}

}  // namespace shrt

extern "C" {

shrt_status shrt_closure_create(int n_grid, int kmax, int width, int layers, int max_batch, uint64_t seed,
                                shrt_closure** out) {
    if (!out) return cfail(SHRT_ERR_INVALID, "closure: null output");
    *out = nullptr;
    if (n_grid < 8 || n_grid > 512 || (n_grid & (n_grid - 1)) || kmax < 1 || 2 * kmax >= n_grid || width < 1 ||
        width > 256 || layers < 1 || layers > 16 || max_batch < 1 || max_batch > 64)
        return cfail(SHRT_ERR_INVALID, "closure: bad shape (n a power of two 8..512, 1 <= kmax < n/2, width <= 256, "
                                       "layers 1..16, max_batch 1..64)");
    shrt_closure* n = new shrt_closure();
    n->N = n_grid; n->kmax = kmax; n->C = width; n->L = layers; n->maxB = max_batch;
    n->NP = (long long)n_grid * n_grid;
    std::vector<int> mkx, mky;
    mode_list(kmax, mkx, mky);
    n->Mm = (int)mkx.size();
    n->D = 2 * n->Mm;
    const int C = width;
    const long long cc = (long long)C * C;
    long long o = 0;
    n->oP = o; o += 4LL * C;
    n->op = o; o += C;
    for (int l = 0; l < layers; ++l) {
        n->oa.push_back(o); o += n->Mm * cc;
        n->ob.push_back(o); o += n->Mm * cc;
        n->oW.push_back(o); o += cc;
        n->ow.push_back(o); o += C;
    }
    n->oQ = o; o += 2LL * C;
    n->oq = o; o += 2;
    n->np = o;

    auto fail_free = [&](shrt_status s) {
        shrt_closure_destroy(n);
        return s;
    };
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        return fail_free(cfail(SHRT_ERR_DEVICE, "closure: stream"));
    n->K1 = kmax + 1;
    n->J = 2 * (2 * kmax + 1);
    n->DP = n->K1 * n->J;
    n->OFF = 2 * kmax;
    const long long NP = n->NP, B = max_batch, D = n->DP;
    shrt_status st = SHRT_OK;
    auto A = [&](float** p, long long count) {
        if (st == SHRT_OK) st = nn_alloc(n, reinterpret_cast<void**>(p), sizeof(float) * (size_t)count);
    };
    A(&n->Fy, (long long)n_grid * 2 * n->K1);
    A(&n->Fxc, (long long)n_grid * n->J);
    A(&n->Fxs, (long long)n_grid * n->J);
    A(&n->T, B * C * n_grid * 2 * n->K1);
    A(&n->prm, n->np); A(&n->grd, n->np); A(&n->am, n->np); A(&n->av, n->np);
    A(&n->x, B * 4 * NP); A(&n->out, B * 2 * NP); A(&n->dout, B * 2 * NP); A(&n->tgt, B * 2 * NP);
    n->h.assign(layers + 1, nullptr);
    n->z.assign(layers, nullptr);
    n->X.assign(layers, nullptr);
    for (int l = 0; l <= layers; ++l) A(&n->h[l], B * C * NP);
    for (int l = 0; l < layers; ++l) { A(&n->z[l], B * C * NP); A(&n->X[l], B * C * D); }
    A(&n->Y, B * C * D); A(&n->dY, B * C * D); A(&n->dX, B * C * D);
    A(&n->dh, B * C * NP); A(&n->dz, B * C * NP);
    A(&n->parts, B * std::max<long long>(cc, 4LL * C));
    n->work_len = std::max<long long>(1LL << 22, 4 * B * C * (long long)D);
    A(&n->work, n->work_len);
    if (st == SHRT_OK) st = nn_alloc(n, reinterpret_cast<void**>(&n->partial), sizeof(double) * n->npartial);
    if (st == SHRT_OK) st = nn_alloc(n, reinterpret_cast<void**>(&n->loss), sizeof(double));
    if (st != SHRT_OK) return fail_free(st);

    // the factors of the orthonormal cos/sin basis (column 2m = cos, 2m+1 = sin
    // of mode m; s = sqrt(2)/N, the (0, 0) mode's 1/N by k_scale_first)
    {
        const int N = n_grid, K1 = n->K1, Jx = n->J;
        std::vector<float> fy((size_t)N * 2 * K1), fxc((size_t)N * Jx), fxs((size_t)N * Jx);
        auto ang = [&](long long k, int i) { return 2.0 * M_PI * (double)(((k * i) % N + N) % N) / N; };
        const double sc = std::sqrt(2.0) / N;
        for (int i = 0; i < N; ++i) {
            for (int ky = 0; ky < K1; ++ky) {
                fy[(size_t)i * 2 * K1 + 2 * ky] = (float)std::cos(ang(ky, i));
                fy[(size_t)i * 2 * K1 + 2 * ky + 1] = (float)std::sin(ang(ky, i));
            }
            for (int kx = -kmax; kx <= kmax; ++kx) {
                const double c = std::cos(ang(kx, i)), sn = std::sin(ang(kx, i));
                const size_t o = (size_t)i * Jx + 2 * (kx + kmax);
                fxc[o] = (float)(sc * c);
                fxc[o + 1] = (float)(sc * sn);
                fxs[o] = (float)(-sc * sn);
                fxs[o + 1] = (float)(sc * c);
            }
        }
        NN_TRY(hipMemcpy(n->Fy, fy.data(), sizeof(float) * fy.size(), hipMemcpyHostToDevice));
        NN_TRY(hipMemcpy(n->Fxc, fxc.data(), sizeof(float) * fxc.size(), hipMemcpyHostToDevice));
        NN_TRY(hipMemcpy(n->Fxs, fxs.data(), sizeof(float) * fxs.size(), hipMemcpyHostToDevice));
        // the non-mode entries of the synthesis inputs stay zero
        NN_TRY(hipMemset(n->Y, 0, sizeof(float) * B * C * D));
        NN_TRY(hipMemset(n->dX, 0, sizeof(float) * B * C * D));
    }
    // parameters: dense weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), spectral
    // weights U(0, 1)/(C*C), biases 0
    {
        std::vector<float> p((size_t)n->np, 0.f);
        uint64_t s = seed;
        auto fill = [&](long long off, long long cnt, double lo, double hi) {
            for (long long i = 0; i < cnt; ++i) p[off + i] = (float)(lo + (hi - lo) * uniform01(s));
        };
        fill(n->oP, 4LL * C, -0.5, 0.5);
        for (int l = 0; l < layers; ++l) {
            fill(n->oa[l], n->Mm * cc, 0.0, 1.0 / cc);
            fill(n->ob[l], n->Mm * cc, 0.0, 1.0 / cc);
            const double r = 1.0 / std::sqrt((double)C);
            fill(n->oW[l], cc, -r, r);
        }
        const double r = 1.0 / std::sqrt((double)C);
        fill(n->oQ, 2LL * C, -r, r);
        NN_TRY(hipMemcpy(n->prm, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
        NN_TRY(hipMemset(n->am, 0, sizeof(float) * n->np));
        NN_TRY(hipMemset(n->av, 0, sizeof(float) * n->np));
    }
    *out = n;
    return SHRT_OK;
}

shrt_status shrt_closure_destroy(shrt_closure* n) {
    if (!n) return SHRT_OK;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    if (n->ev_in) (void)hipEventDestroy(n->ev_in);
    if (n->ev_out) (void)hipEventDestroy(n->ev_out);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return SHRT_OK;
}

shrt_status shrt_closure_info(shrt_closure* n, long long* nparams, int* nmodes) {
    if (!n) return cfail(SHRT_ERR_INVALID, "closure: null handle");
    if (nparams) *nparams = n->np;
    if (nmodes) *nmodes = n->Mm;
    return SHRT_OK;
}

shrt_status shrt_closure_get_params(shrt_closure* n, float* host) {
    if (!n || !host) return cfail(SHRT_ERR_INVALID, "closure: null argument");
    NN_TRY(hipMemcpyAsync(host, n->prm, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    NN_TRY(hipStreamSynchronize(n->stream));
    return SHRT_OK;
}

shrt_status shrt_closure_set_params(shrt_closure* n, const float* host) {
    if (!n || !host) return cfail(SHRT_ERR_INVALID, "closure: null argument");
    NN_TRY(hipMemcpyAsync(n->prm, host, sizeof(float) * n->np, hipMemcpyHostToDevice, n->stream));
    NN_TRY(hipMemsetAsync(n->am, 0, sizeof(float) * n->np, n->stream));
    NN_TRY(hipMemsetAsync(n->av, 0, sizeof(float) * n->np, n->stream));
    NN_TRY(hipStreamSynchronize(n->stream));
    n->adam_t = 0;
    return SHRT_OK;
}

shrt_status shrt_closure_forward(shrt_closure* n, const float* in, int batch, float* out) {
    if (!n || !in || !out || batch < 1 || batch > n->maxB) return cfail(SHRT_ERR_INVALID, "closure: bad arguments");
    NN_TRY(hipMemcpyAsync(n->x, in, sizeof(float) * batch * 4 * n->NP, hipMemcpyHostToDevice, n->stream));
    NN_TRY(forward(n, batch));
    NN_TRY(hipMemcpyAsync(out, n->out, sizeof(float) * batch * 2 * n->NP, hipMemcpyDeviceToHost, n->stream));
    NN_TRY(hipStreamSynchronize(n->stream));
    return SHRT_OK;
}

shrt_status shrt_closure_gradients(shrt_closure* n, const float* in, const float* target, int batch, double* loss,
                                   float* grad) {
    if (!n || !in || !target || batch < 1 || batch > n->maxB) return cfail(SHRT_ERR_INVALID, "closure: bad arguments");
    NN_TRY(hipMemcpyAsync(n->x, in, sizeof(float) * batch * 4 * n->NP, hipMemcpyHostToDevice, n->stream));
    NN_TRY(hipMemcpyAsync(n->tgt, target, sizeof(float) * batch * 2 * n->NP, hipMemcpyHostToDevice, n->stream));
    NN_TRY(forward(n, batch));
    NN_TRY(loss_and_grad(n, batch));
    NN_TRY(backward(n, batch));
    if (loss) NN_TRY(hipMemcpyAsync(loss, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
    if (grad) NN_TRY(hipMemcpyAsync(grad, n->grd, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream));
    NN_TRY(hipStreamSynchronize(n->stream));
    return SHRT_OK;
}

shrt_status shrt_closure_set_data(shrt_closure* n, const float* in, const float* target, int nsamples) {
    if (!n || !in || !target || nsamples < 1) return cfail(SHRT_ERR_INVALID, "closure: bad arguments");
    if (n->din) { (void)hipFree(n->din); n->din = nullptr; }
    if (n->dtg) { (void)hipFree(n->dtg); n->dtg = nullptr; }
    if (n->idx) { (void)hipFree(n->idx); n->idx = nullptr; }
    n->nsamples = 0;
    if (hipMalloc(&n->din, sizeof(float) * (size_t)nsamples * 4 * n->NP) != hipSuccess ||
        hipMalloc(&n->dtg, sizeof(float) * (size_t)nsamples * 2 * n->NP) != hipSuccess ||
        hipMalloc(&n->idx, sizeof(int) * n->maxB) != hipSuccess)
        return cfail(SHRT_ERR_ALLOC, "closure: data set allocation failed");
    NN_TRY(hipMemcpy(n->din, in, sizeof(float) * (size_t)nsamples * 4 * n->NP, hipMemcpyHostToDevice));
    NN_TRY(hipMemcpy(n->dtg, target, sizeof(float) * (size_t)nsamples * 2 * n->NP, hipMemcpyHostToDevice));
    n->nsamples = nsamples;
    return SHRT_OK;
}

shrt_status shrt_closure_train(shrt_closure* n, int iters, int batch, float lr, uint64_t seed, double* losses) {
    if (!n || iters < 1 || batch < 1 || batch > n->maxB || !(lr > 0.f)) return cfail(SHRT_ERR_INVALID, "closure: bad arguments");
    if (n->nsamples < 1) return cfail(SHRT_ERR_INVALID, "closure: no training data (shrt_closure_set_data)");
    uint64_t s = seed;
    std::vector<int> pick(batch);
    std::vector<double> loss_h(iters);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; ++it) {
        for (int b = 0; b < batch; ++b) pick[b] = (int)(splitmix(s) % (uint64_t)n->nsamples);
        NN_TRY(hipMemcpyAsync(n->idx, pick.data(), sizeof(int) * batch, hipMemcpyHostToDevice, n->stream));
        hipLaunchKernelGGL(k_gather, pw_grid(batch * 4 * n->NP), dim3(256), 0, n->stream, n->din, n->idx, n->x, batch,
                           4 * n->NP);
        hipLaunchKernelGGL(k_gather, pw_grid(batch * 2 * n->NP), dim3(256), 0, n->stream, n->dtg, n->idx, n->tgt, batch,
                           2 * n->NP);
        NN_TRY(forward(n, batch));
        NN_TRY(loss_and_grad(n, batch));
        NN_TRY(backward(n, batch));
        n->adam_t += 1;
        const double t = (double)n->adam_t;
        const float lr_t = (float)(lr * std::sqrt(1.0 - std::pow(b2, t)) / (1.0 - std::pow(b1, t)));
        hipLaunchKernelGGL(k_adam, pw_grid(n->np), dim3(256), 0, n->stream, n->prm, n->grd, n->am, n->av, n->np, lr_t,
                           b1, b2, eps);
        NN_TRY(hipGetLastError());
        NN_TRY(hipMemcpyAsync(&loss_h[it], n->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
        // the pick buffer is rewritten next iteration: wait for this one's gather
        NN_TRY(hipStreamSynchronize(n->stream));
    }
    if (losses) std::memcpy(losses, loss_h.data(), sizeof(double) * iters);
    return SHRT_OK;
}

}  // extern "C"

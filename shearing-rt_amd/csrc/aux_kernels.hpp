// aux_kernels.hpp -- transforms and reductions outside the step:
//   xtok (fft.cu:136-166)   = k_aux_xfwd then k_aux_yfwd
//   ktox (fft.cu:170-204)   = k_aux_yinv then k_aux_xinv, with optional ik
//                             multipliers for the velocities (shear.cu:378-391)
//   shearing_field view     (shear.cu:296-322)
//   neg_lapinv_shear        (shear.cu:367-374) for aphi after initialisation
//   FP64 compensated energy / enstrophy / density-variance reductions (new)
// They reuse the team FFT and the G/H tile layout of the step kernels.
#pragma once

#include "step_common.hpp"

namespace shrt {

template <typename T>
struct AuxArgs {
    Geo g;
    const void* in;
    void* out;
    const double* kx;
    const double* ky;
    const double* s;      // shift table for the multipliers
    const int* jmp;       // remap view of `in` (spectral inputs), or null
    const cplx<T>* tw;
    int mode;             // k_aux_yinv: 0 identity, 1 i*kys*F (vx of phi), 2 -i*kx*F (vy of phi)
};

// x-forward of one real field f[ix*ny + iy] for a tile of 4 y columns -> H tile.
template <typename T, int NX>
__global__ void __launch_bounds__((kTile * TeamShape<T, NX>::TT))
k_aux_xfwd(AuxArgs<T> a) {
    using S = TeamShape<T, NX>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NX>();
    constexpr int NQ = (NX / 3 + 1 + TT - 1) / TT;
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, y = tile * kTile + team, nk = g.nkxh;
    const T* f = static_cast<const T*>(a.in);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = {f[(long long)(tid + m * TT) * g.ny + y], T(0)};
    TeamFFT<T, NX, P, S::RMAX, -1>::run(v, lds, a.tw, tid, sync);
#pragma unroll
    for (int m = 0; m < P; ++m) lds[tid + m * TT] = v[m];
    __syncthreads();
    cplx<T> h[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int kx = tid + i * TT;
        if (kx < nk) h[i] = lds[kx];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int kx = tid + i * TT;
        if (kx < nk) sm[kx * 4 + team] = h[i];
    }
    __syncthreads();
    cplx<T>* Ho = static_cast<cplx<T>*>(a.out) + (long long)tile * nk * 4;
    for (int idx = threadIdx.x; idx < nk * 4; idx += blockDim.x) Ho[idx] = sm[idx];
}

// y-forward of an H field for a tile of kx>=0 rows -> F (reference layout),
// truncated and scaled (scale_dealias).
template <typename T, int NY>
__global__ void __launch_bounds__((kTile * TeamShape<T, NY>::TT))
k_aux_yfwd(AuxArgs<T> a) {
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, row = tile * kTile + team;
    const bool act = row < g.nkxh;
    const int rn = row == 0 ? 0 : g.nkx - row;
    const int nrow_tile = min(kTile, g.nkxh - tile * kTile);
    const cplx<T>* Hf = static_cast<const cplx<T>*>(a.in);
    cplx<T>* F = static_cast<cplx<T>*>(a.out);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    constexpr int LP = stage_stride<NY>();
    for (int idx = threadIdx.x; idx < NY * 4; idx += blockDim.x) {
        const int t = idx >> 4, w = idx & 15;
        sm[(w >> 2) * LP + 4 * t + (w & 3)] =
            (w >> 2) < nrow_tile ? Hf[((long long)t * g.nkxh + tile * kTile) * 4 + w] : czero<T>();
    }
    __syncthreads();
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = sm[team * LP + tid + m * TT];
    __syncthreads();
    TeamFFT<T, NY, P, S::RMAX, -1>::run(v, lds, a.tw, tid, sync);
#pragma unroll
    for (int m = 0; m < P; ++m) lds[tid + m * TT] = v[m];
    __syncthreads();
    if (!act) return;
    const T nn = (T)((double)g.nx * (double)g.ny);   // scale_dealias divides by nx*ny
    for (int iky = tid; iky < g.nky; iky += TT) {
        const cplx<T> p = lds[iky];
        F[(long long)row * g.nky + iky] = {p.x / nn, p.y / nn};
        if (row > 0) {
            const cplx<T> q = lds[(NY - iky) & (NY - 1)];
            F[(long long)rn * g.nky + iky] = {q.x / nn, -q.y / nn};
        }
    }
}

// y-inverse of a spectral field (reference layout, through the remap view),
// times the multiplier of `mode`, for a tile of kx>=0 rows -> G tile.
template <typename T, int NY>
__global__ void __launch_bounds__((kTile * TeamShape<T, NY>::TT))
k_aux_yinv(AuxArgs<T> a) {
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, row = tile * kTile + team;
    const bool act = row < g.nkxh;
    const int rn = row == 0 ? 0 : g.nkx - row;
    const int nrow_tile = min(kTile, g.nkxh - tile * kTile);
    const cplx<T>* F = static_cast<const cplx<T>*>(a.in);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    auto mult = [&](int r, int iky, cplx<T> z) -> cplx<T> {
        if (a.mode == 1) {
            const T k = (T)kys_of(a.ky, a.s, r, iky);
            return {-k * z.y, k * z.x};
        }
        if (a.mode == 2) {
            const T k = (T)a.kx[r];
            return {k * z.y, -k * z.x};
        }
        return z;
    };
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        cplx<T> val = czero<T>();
        if (act) {
            if (e < g.nky) {
                val = mult(row, e, view_load(F, row, e, g, a.jmp));
                if (e == 0) {
                    const cplx<T> vn = mult(rn, 0, view_load(F, rn, 0, g, a.jmp));
                    val = {(val.x + vn.x) * T(0.5), (val.y - vn.y) * T(0.5)};
                }
            } else if (e > NY - g.nky) {
                const int iky = NY - e;
                const cplx<T> vn = mult(rn, iky, view_load(F, rn, iky, g, a.jmp));
                val = {vn.x, -vn.y};
            }
        }
        v[m] = val;
    }
    TeamFFT<T, NY, P, S::RMAX, +1>::run(v, lds, a.tw, tid, sync);
    constexpr int LP = stage_stride<NY>();
#pragma unroll
    for (int m = 0; m < P; ++m) sm[team * LP + tid + m * TT] = v[m];
    __syncthreads();
    cplx<T>* Gf = static_cast<cplx<T>*>(a.out);
    for (int idx = threadIdx.x; idx < NY * 4; idx += blockDim.x) {
        const int t = idx >> 4, w = idx & 15;
        if ((w >> 2) < nrow_tile)
            Gf[((long long)t * g.nkxh + tile * kTile) * 4 + w] = sm[(w >> 2) * LP + 4 * t + (w & 3)];
    }
}

// x-inverse (C2R) of a G field for a tile of 4 y columns -> real f[ix*ny + iy].
template <typename T, int NX>
__global__ void __launch_bounds__((kTile * TeamShape<T, NX>::TT))
k_aux_xinv(AuxArgs<T> a) {
    using S = TeamShape<T, NX>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NX>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, y = tile * kTile + team, nk = g.nkxh;
    const cplx<T>* Gf = static_cast<const cplx<T>*>(a.in) + (long long)tile * nk * 4;
    T* f = static_cast<T*>(a.out);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    for (int idx = threadIdx.x; idx < nk * 4; idx += blockDim.x) sm[idx] = Gf[idx];
    __syncthreads();
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        cplx<T> z = czero<T>();
        if (e < nk) {
            const cplx<T> A = sm[e * 4 + team];
            z = e == 0 ? cplx<T>{A.x, T(0)} : A;
        } else if (e > NX - nk) {
            z = cconj(sm[(NX - e) * 4 + team]);
        }
        v[m] = z;
    }
    __syncthreads();
    TeamFFT<T, NX, P, S::RMAX, +1>::run(v, lds, a.tw, tid, sync);
#pragma unroll
    for (int m = 0; m < P; ++m) f[(long long)(tid + m * TT) * g.ny + y] = v[m].x;
}

// Lab-frame output transform ktox_shear (shear.cu:467-478), y part
// (idft_shear_y, :481-513): for every stored kx row (both signs) the one-sided
// ky sum with weights 1 (ky = 0) and 2 (ky > 0) at the sheared wavenumber
// ky + ky_shift[kx].  Here it is an inverse FFT of length NY of the weighted
// row (zero for ky >= nky) followed by the phase exp(i ky_shift[kx] yy[iy]),
// yy[iy] = iy*dy, formed in FP64.  Output tiles [t][row][c] (y = 4t + c), row
// stride nkx.
template <typename T, int NY>
__global__ void __launch_bounds__((kTile * TeamShape<T, NY>::TT))
k_aux_yinv_lab(AuxArgs<T> a) {
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    constexpr int LP = stage_stride<NY>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, row = tile * kTile + team;
    const bool act = row < g.nkx;
    const int nrow_tile = min(kTile, g.nkx - tile * kTile);
    const cplx<T>* F = static_cast<const cplx<T>*>(a.in);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        cplx<T> z = czero<T>();
        if (act && e < g.nky) {
            z = view_load(F, row, e, g, a.jmp);
            if (e > 0) z = {T(2) * z.x, T(2) * z.y};
        }
        v[m] = z;
    }
    TeamFFT<T, NY, P, S::RMAX, +1>::run(v, lds, a.tw, tid, sync);
    const double sh = act ? a.s[row] : 0.0;
#pragma unroll
    for (int m = 0; m < P; ++m) {
        double sn, cs;
        sincos(mul_rn(sh, (double)(tid + m * TT) * g.dy), &sn, &cs);
        v[m] = cmul(v[m], cplx<T>{(T)cs, (T)sn});
    }
#pragma unroll
    for (int m = 0; m < P; ++m) sm[team * LP + tid + m * TT] = v[m];
    __syncthreads();
    cplx<T>* Gl = static_cast<cplx<T>*>(a.out);
    for (int idx = threadIdx.x; idx < NY * 4; idx += blockDim.x) {
        const int t = idx >> 4, w = idx & 15;
        if ((w >> 2) < nrow_tile)
            Gl[((long long)t * g.nkx + tile * kTile) * 4 + w] = sm[(w >> 2) * LP + 4 * t + (w & 3)];
    }
}

// Lab-frame output transform, x part (pad2d, shear.cu:515-527, and ktox_1d,
// fft.cu:208-250): per y column, the stored kx rows padded to NX, a complex
// inverse FFT, and the real part -> f[ix*ny + iy].
template <typename T, int NX>
__global__ void __launch_bounds__((kTile * TeamShape<T, NX>::TT))
k_aux_xinv_lab(AuxArgs<T> a) {
    using S = TeamShape<T, NX>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NX>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int tile = blockIdx.x, y = tile * kTile + team, nk = g.nkx;
    const cplx<T>* Gl = static_cast<const cplx<T>*>(a.in) + (long long)tile * nk * 4;
    T* f = static_cast<T*>(a.out);
    cplx<T>* lds = sm + team * L;
    auto sync = [] { __syncthreads(); };
    for (int idx = threadIdx.x; idx < nk * 4; idx += blockDim.x) sm[(idx & 3) * nk + (idx >> 2)] = Gl[idx];
    __syncthreads();
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        const int i = e < g.nkxh ? e : (e > g.nkxh2 ? e - g.nkxpad : -1);
        v[m] = i >= 0 ? sm[team * nk + i] : czero<T>();
    }
    __syncthreads();
    TeamFFT<T, NX, P, S::RMAX, +1>::run(v, lds, a.tw, tid, sync);
#pragma unroll
    for (int m = 0; m < P; ++m) f[(long long)(tid + m * TT) * g.ny + y] = v[m].x;
}

// out[t] = remap view of in (shear.cu:296-322); jmp null -> plain copy.
template <typename T>
__global__ void k_view_copy(Geo g, const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out,
                            const int* __restrict__ jmp) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nkx * g.nky) return;
    const int row = (int)(t / g.nky), iky = (int)(t - (long long)row * g.nky);
    out[t] = view_load(in, row, iky, g, jmp);
}

// aphi = omega / kperp2_shear for modes 1..M-1, aphi[0] = 0 (fourier.cu:233-240)
// Modes with |kx| > kcut or ky > kcut (integer wavenumbers, storage order of
// fourier.cu:124-125) set to zero.
template <typename T>
__global__ void k_lowpass(Geo g, cplx<T>* __restrict__ f, int kcut) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nkx * g.nky) return;
    const int row = (int)(t / g.nky), iky = (int)(t - (long long)row * g.nky);
    const int kx = row < g.nkxh ? row : row - g.nkx;
    if (kx > kcut || -kx > kcut || iky > kcut) f[t] = czero<T>();
}

template <typename T>
__global__ void k_phi_from_omega(Geo g, const cplx<T>* __restrict__ om, cplx<T>* __restrict__ phi,
                                 const double* kx, const double* ky, const double* s) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nkx * g.nky) return;
    const int row = (int)(t / g.nky), iky = (int)(t - (long long)row * g.nky);
    if (t == 0) { phi[0] = czero<T>(); return; }
    const T k2 = (T)k2_of(kx[row], kys_of(ky, s, row, iky));
    const cplx<T> w = om[t];
    phi[t] = {w.x / k2, w.y / k2};
}

// Plain 16-byte copy: the reference point for HBM traffic counters
// (reads and writes exactly `n` 16-byte words).
template <typename T>
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// Closure network input channel from a physical field: x = (float)(f / scale)
// with the division in FP64 (what the host pipeline computes before its FP32
// conversion); scale = 0 fills the channel with the constant `fill`.
template <typename T>
__global__ void k_net_in(const T* __restrict__ f, float* __restrict__ x, long long n, double scale, double fill) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    x[t] = scale != 0.0 ? (float)((double)f[t] / scale) : (float)fill;
}

// Physical correction field from a network output channel: f = (T)(out * scale).
template <typename T>
__global__ void k_net_out(const float* __restrict__ o, T* __restrict__ f, long long n, double scale) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = (T)((double)o[t] * scale);
}

// One workgroup of 1024 threads: out[0] = sum |a|^2 over n complex values in
// FP64, or (nonfinite) out[0] = the number of non-finite values among n floats.
template <typename T>
__global__ void __launch_bounds__(1024) k_sumsq(const cplx<T>* __restrict__ a, const float* __restrict__ o,
                                                 long long n, double* __restrict__ out) {
    __shared__ double red[1024];
    double s = 0.0;
    for (long long t = threadIdx.x; t < n; t += blockDim.x) {
        if (a) {
            const double x = (double)a[t].x, y = (double)a[t].y;
            s += x * x + y * y;
        } else {
            s += isfinite(o[t]) ? 0.0 : 1.0;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// dst = a + (T)dt * c per component, each operation rounded in T (no
// contraction): the host pipeline's state + dt * correction in the state's type.
template <typename T>
__global__ void k_axpy_round(const cplx<T>* __restrict__ a, const cplx<T>* __restrict__ c, double dt,
                             cplx<T>* __restrict__ dst, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if constexpr (sizeof(T) == 4) {
        const float d = (float)dt;
        dst[t] = {add_rn(a[t].x, mul_rn(d, c[t].x)), add_rn(a[t].y, mul_rn(d, c[t].y))};
    } else {
        dst[t] = {add_rn(a[t].x, mul_rn(dt, c[t].x)), add_rn(a[t].y, mul_rn(dt, c[t].y))};
    }
}

// The fastest copy measured on MI355X (the achievable HBM rate): four 16-byte
// words per thread in flight, nontemporal loads and stores (whole lines, so
// nothing is gained by keeping them in the caches).
template <typename T>
__global__ void k_copy_stream(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n) {
    using W = unsigned int __attribute__((ext_vector_type(4)));
    const W* s = reinterpret_cast<const W*>(src);
    W* d = reinterpret_cast<W*>(dst);
    const long long base = (long long)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    W v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = base + (long long)u * blockDim.x;
        v[u] = i < n ? __builtin_nontemporal_load(s + i) : W{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = base + (long long)u * blockDim.x;
        if (i < n) __builtin_nontemporal_store(v[u], d + i);
    }
}

template <typename T>
__global__ void k_fill_zero(cplx<T>* p, long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = czero<T>();
}

// out = (accumulate ? out : 0) + (negate ? -a*b : a*b), pointwise on physical
// fields: the product kernel of poisson_bracket_shear (shear.cu:415-426), one
// product per launch.
template <typename T>
__global__ void k_mul_acc(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out, long long n,
                          int negate, int accumulate) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const T p = negate ? -a[t] : a[t];
    out[t] = accumulate ? fma(p, b[t], out[t]) : p * b[t];
}

// Neumaier-compensated FP64 accumulation.
struct KSum {
    double s, c;
    __device__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ void merge(KSum o) { add(o.s); add(o.c); }
};

constexpr int kDiagThreads = 256;
constexpr int kDiagQ = 3;   // energy, enstrophy, density variance

// Per-block partial sums of E = 1/2 sum w kperp2 |phi|^2, Z = 1/2 sum w |omega|^2,
// R = 1/2 sum w |rho|^2 over the stored half plane (w = 1 for iky = 0, else 2).
template <typename T>
__global__ void __launch_bounds__(kDiagThreads)
k_diag_partial(Geo g, const cplx<T>* __restrict__ om, const cplx<T>* __restrict__ rh, const int* jmp,
               const double* kx, const double* ky, const double* s, double* partial) {
    __shared__ double sh[kDiagQ * 2][kDiagThreads];
    KSum acc[kDiagQ];
#pragma unroll
    for (int q = 0; q < kDiagQ; ++q) acc[q] = {0.0, 0.0};
    const long long M = (long long)g.nkx * g.nky;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < M;
         t += (long long)gridDim.x * blockDim.x) {
        const int row = (int)(t / g.nky), iky = (int)(t - (long long)row * g.nky);
        const double w = iky == 0 ? 1.0 : 2.0;
        const cplx<T> o = view_load(om, row, iky, g, jmp);
        const cplx<T> r = view_load(rh, row, iky, g, jmp);
        const double ox = o.x, oy = o.y, rx = r.x, ry = r.y;
        const double o2 = ox * ox + oy * oy;
        if (t > 0) {
            const double k2 = k2_of(kx[row], kys_of(ky, s, row, iky));
            acc[0].add(w * o2 / k2);   // kperp2 |omega/kperp2|^2
        }
        acc[1].add(w * o2);
        acc[2].add(w * (rx * rx + ry * ry));
    }
#pragma unroll
    for (int q = 0; q < kDiagQ; ++q) {
        sh[2 * q][threadIdx.x] = acc[q].s;
        sh[2 * q + 1][threadIdx.x] = acc[q].c;
    }
    __syncthreads();
    if (threadIdx.x < kDiagQ) {
        const int q = threadIdx.x;
        KSum tot = {0.0, 0.0};
        for (int i = 0; i < kDiagThreads; ++i) tot.merge({sh[2 * q][i], sh[2 * q + 1][i]});
        partial[(long long)blockIdx.x * kDiagQ * 2 + 2 * q] = tot.s;
        partial[(long long)blockIdx.x * kDiagQ * 2 + 2 * q + 1] = tot.c;
    }
}

template <typename T>
__global__ void k_diag_final(const double* partial, int nblocks, double* out) {
    if (threadIdx.x >= kDiagQ) return;
    const int q = threadIdx.x;
    KSum tot = {0.0, 0.0};
    for (int b = 0; b < nblocks; ++b) tot.merge({partial[b * kDiagQ * 2 + 2 * q], partial[b * kDiagQ * 2 + 2 * q + 1]});
    out[q] = 0.5 * (tot.s + tot.c);
}

}  // namespace shrt

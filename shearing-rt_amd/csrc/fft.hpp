// fft.hpp -- LDS-resident Stockham FFT for CDNA4 (gfx950), used by every
// transform of the pseudospectral step.
//
// A "team" of TT = N/P threads transforms one length-N complex sequence.  Thread
// t holds the P elements t + m*TT (m = 0..P-1) in registers, in natural order,
// both on entry and on exit.  Each stage runs P/R radix-R butterflies per thread
// in registers (R <= RMAX), writes its outputs to the team's LDS buffer at the
// auto-sorted Stockham positions and reads its next inputs back in natural
// order; the last stage needs no exchange.  Inter-stage twiddles come from a
// per-length table W_N^k (rounded once from long double on the host), so the
// transform is as accurate as the table and the radix kernels.
//
// Replaces the cuFFT 2-D R2C/C2R plans of fft.cu:111-119, 147-149, 183-185.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

namespace shrt {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T> __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }
template <typename T> __device__ __forceinline__ cplx<T> cscale(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
// a * (i*d), d = +1 or -1
template <int D, typename T> __device__ __forceinline__ cplx<T> cmul_i(cplx<T> a) {
    if constexpr (D > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

// Separately rounded arithmetic that is never fused into an FMA.  The
// toolchain defines __dadd_rn, __dmul_rn and the rest of that family as plain
// operators, which -ffp-contract=fast contracts (a*b + c becomes one FMA);
// these keep every operation rounded as written, for the FP64 shift
// recurrence, kperp2, the CFL tests and the other places that follow the
// reference's unfused arithmetic.
#define SHRT_RN_OP(name, op)                                             \
    __device__ __forceinline__ double name(double a, double b) {         \
        _Pragma("clang fp contract(off)") return a op b;                 \
    }                                                                    \
    __device__ __forceinline__ float name(float a, float b) {            \
        _Pragma("clang fp contract(off)") return a op b;                 \
    }
SHRT_RN_OP(add_rn, +)
SHRT_RN_OP(sub_rn, -)
SHRT_RN_OP(mul_rn, *)
SHRT_RN_OP(div_rn, /)
#undef SHRT_RN_OP

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// cos(2*pi*j/32) for j = 0..8; other angles follow by symmetry.
__host__ __device__ constexpr double cos32_q(int j) {
    return j == 0 ? 1.0
         : j == 1 ? 0.98078528040323044912618
         : j == 2 ? 0.92387953251128675612818
         : j == 3 ? 0.83146961230254523707879
         : j == 4 ? 0.70710678118654752440084
         : j == 5 ? 0.55557023301960222474283
         : j == 6 ? 0.38268343236508977172846
         : j == 7 ? 0.19509032201612826784828
         : 0.0;
}
__host__ __device__ constexpr double cos32(int j) {
    j = ((j % 32) + 32) % 32;
    return j <= 8 ? cos32_q(j) : j <= 16 ? -cos32_q(16 - j) : j <= 24 ? -cos32_q(j - 16) : cos32_q(32 - j);
}
__host__ __device__ constexpr double sin32(int j) { return cos32(8 - j); }

// v * exp(D * 2*pi*i * M / R), D = -1 forward, +1 inverse; R divides 32.
template <int R, int M, int D, typename T>
__device__ __forceinline__ cplx<T> twc(cplx<T> v) {
    constexpr int J = ((M * (32 / R)) % 32 + 32) % 32;
    if constexpr (J == 0) return v;
    else if constexpr (J == 16) return {-v.x, -v.y};
    else if constexpr (J == 8) return cmul_i<D>(v);
    else if constexpr (J == 24) return cmul_i<-D>(v);
    else {
        constexpr T c = (T)cos32(J);
        constexpr T s = (T)(D * sin32(J));
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

// In-register radix-R DFT, natural order in and out.  R = R1*R2 composites are
// split as n = R2*n1 + n2, k = k1 + R1*k2 (decimation in time).
template <int R, int D, typename T>
__device__ __forceinline__ void dft(cplx<T>* v) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cplx<T> t2 = cadd(v[1], v[3]), t3 = cmul_i<D>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else {
        constexpr int R1 = 4, R2 = R / 4;
        static_assert(R % 4 == 0, "radix must be 2, 4 or a multiple of 4");
        cplx<T> y[R];
        static_for<0, R2>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            cplx<T> t[R1];
            static_for<0, R1>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                t[n1] = v[R2 * n1 + n2];
            });
            dft<R1, D>(t);
            static_for<0, R1>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                y[k1 * R2 + n2] = twc<R, n2 * k1, D>(t[k1]);
            });
        });
        static_for<0, R1>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            cplx<T> t[R2];
            static_for<0, R2>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                t[n2] = y[k1 * R2 + n2];
            });
            dft<R2, D>(t);
            static_for<0, R2>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                v[k1 + R1 * k2] = t[k2];
            });
        });
    }
}

// LDS index with one pad slot every 32 elements (spreads the strided Stockham
// stores over the 64 LDS banks).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
template <int N> constexpr int lds_len() { return N + N / 32; }

// Radix and butterflies per thread of the stage that starts at sub-length NS.
template <int N, int P, int RMAX, int NS>
struct StageShape {
    static constexpr int REM = N / NS;
    static constexpr int R0 = REM < RMAX ? REM : RMAX;
    static constexpr int R = R0 < P ? R0 : P;
    static constexpr int Q = P / R;
    static constexpr bool LAST = (NS * R == N);
};

// The table twiddles of every stage of one team transform, held in registers:
// w_1 (and w_4 for radix >= 8) per butterfly, W_{NS*R}^{k} = tw[k*N/(NS*R)] for
// the butterfly's position k.  Loading them before a transform keeps global
// loads out of the transform, so loads a caller issues ahead of it stay in
// flight (vector-memory loads complete in order).  The same values serve the
// forward and the inverse transform.
template <typename T, int N, int P, int RMAX, int NS = 1, bool MORE = (NS < N)>
struct TwiddleSet {
    using Sh = StageShape<N, P, RMAX, NS>;
    static constexpr int TT = N / P;
    cplx<T> w1[Sh::Q], w4[Sh::Q];
    TwiddleSet<T, N, P, RMAX, NS * Sh::R> next;

    __device__ __forceinline__ void load(const cplx<T>* __restrict__ tw, int tid) {
        if constexpr (NS > 1) {
            constexpr int UNIT = N / (NS * Sh::R);
#pragma unroll
            for (int q = 0; q < Sh::Q; ++q) {
                const int k = (tid + q * TT) & (NS - 1);
                w1[q] = tw[k * UNIT];
                if constexpr (Sh::R >= 8) w4[q] = tw[4 * k * UNIT];
            }
        }
        next.load(tw, tid);
    }
};
template <typename T, int N, int P, int RMAX, int NS>
struct TwiddleSet<T, N, P, RMAX, NS, false> {
    __device__ __forceinline__ void load(const cplx<T>*, int) {}
};

// Team FFT of length N, P points per thread (TT = N/P threads), radix <= RMAX.
// D = -1 forward (e^{-i}), +1 inverse; unnormalised.  `tw` holds W_N^k = e^{-2 pi i k/N}.
// `lds` is the team's private buffer of lds_len<N>() elements.  SYNC is called
// between the LDS store and load of every exchange and after the load.
template <typename T, int N, int P, int RMAX, int D>
struct TeamFFT {
    static constexpr int TT = N / P;
    static_assert(N % P == 0, "P must divide N");
    using Twiddles = TwiddleSet<T, N, P, RMAX>;

    template <int NS, typename TS, typename SYNC>
    __device__ static __forceinline__ void stages(cplx<T>* v, cplx<T>* lds, const TS& ts, int tid, SYNC&& sync) {
        using Sh = StageShape<N, P, RMAX, NS>;
        constexpr int R = Sh::R, Q = Sh::Q;
        constexpr bool LAST = Sh::LAST;
        static_assert(P % R == 0, "stage radix must divide P");
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            cplx<T> u[R];
            const int b = tid + q * TT;
            const int k = b & (NS - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
            if constexpr (NS > 1) {
                // w_r = W_{NS*R}^{r*k}: w_1 (and w_4 for R >= 8) from the table,
                // the others as short products of those (relative error <= ~8 ulp).
                // Each power is applied as soon as it is formed; only w_1..w_4
                // and the multiples of 4 stay live.
                cplx<T> w1 = ts.w1[q];
                if constexpr (D > 0) w1.y = -w1.y;
                u[1] = cmul(u[1], w1);
                if constexpr (R >= 4) {
                    const cplx<T> w2 = cmul(w1, w1);
                    const cplx<T> w3 = cmul(w2, w1);
                    u[2] = cmul(u[2], w2);
                    u[3] = cmul(u[3], w3);
                    if constexpr (R >= 8) {
                        cplx<T> w4 = ts.w4[q];
                        if constexpr (D > 0) w4.y = -w4.y;
                        cplx<T> w4m = w4;   // w_{4j}
#pragma unroll
                        for (int j = 1; j < R / 4; ++j) {
                            if (j > 1) w4m = cmul(w4m, w4);
                            u[4 * j] = cmul(u[4 * j], w4m);
                            u[4 * j + 1] = cmul(u[4 * j + 1], cmul(w4m, w1));
                            u[4 * j + 2] = cmul(u[4 * j + 2], cmul(w4m, w2));
                            u[4 * j + 3] = cmul(u[4 * j + 3], cmul(w4m, w3));
                        }
                    }
                }
            }
            dft<R, D>(u);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[q + r * Q] = u[r];
            } else {
                // lpad(B0 + k + r*NS) = [B0 + k + (B0 >> 5) + (k >> 5)] + [r*NS + (r*NS >> 5)]
                // for the block base B0 = (b/NS)*NS*R (a multiple of 32, or a block
                // inside one 32-element segment) and k < NS: the second bracket is a
                // compile-time constant, so the stores use immediate offsets.
                const int B0 = (b / NS) * NS * R;
                cplx<T>* dst = lds + (B0 + k + (B0 >> 5) + (k >> 5));
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * NS + ((r * NS) >> 5)] = u[r];
            }
        }
        if constexpr (!LAST) {
            sync();
            if constexpr (TT % 32 == 0) {
                const cplx<T>* src = lds + (tid + (tid >> 5));
#pragma unroll
                for (int m = 0; m < P; ++m) v[m] = src[m * (TT + TT / 32)];
            } else {
#pragma unroll
                for (int m = 0; m < P; ++m) v[m] = lds[lpad(tid + m * TT)];
            }
            sync();
            stages<NS * R>(v, lds, ts.next, tid, sync);
        }
    }

    // Transform with twiddles loaded by the caller beforehand.
    template <typename SYNC>
    __device__ static __forceinline__ void run(cplx<T>* v, cplx<T>* lds, const Twiddles& ts, int tid, SYNC&& sync) {
        stages<1>(v, lds, ts, tid, sync);
    }

    template <typename SYNC>
    __device__ static __forceinline__ void run(cplx<T>* v, cplx<T>* lds, const cplx<T>* __restrict__ tw, int tid,
                                               SYNC&& sync) {
        Twiddles ts;
        ts.load(tw, tid);
        stages<1>(v, lds, ts, tid, sync);
    }
};

// Ordering of a team's LDS exchange when the team is one wavefront: the lanes
// of a wave issue their LDS operations in program order, so only the compiler
// must be kept from moving loads across the stores.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Team geometry with an explicit point count per thread.
template <int N, int PTS>
struct Team {
    static constexpr int P = PTS;
    static constexpr int TT = N / PTS;
    static constexpr int RMAX = 16;
};

// Team geometry for a transform length in precision T: P points per thread,
// TT threads.  16 points per thread from N = 2048 up in FP32 (2 wavefronts per
// team at 2048, 4 at 4096) and at 4096 in FP64; 4 at 1024 in FP64; 8 below.
// Measured per launch on MI355X: 1024^2 FP32 step 37.7 us with 8 points
// against 45.2 us with 16, FP64 68.9 -> 55.9 us, and FP64 with 4 points 50.1
// against 51.2 us (4 waves per team: P3 14.5 against 15.1), while FP32 with 4
// points took 40.5 against 35.3 (the fused P3+P1 28.7 against 23.4); 2048^2 FP32 with 8
// points P1 29.9 against 25.5 us, P3 33.4 against 30.0 (16 kept); 2048^2 FP64
// with 8 points P1 40.6 against 44.5 us, P3 45.1 against 47.1 (8 taken; P2
// keeps 16, p2_points), with 4 points 164.4 against 156.9 us per step (P3 58.8
// against 51.2); 4096^2 FP32 with 8 points 385-398 against 369 us per step (16
// kept).
template <typename T, int N>
struct TeamShape {
    static constexpr int P = (sizeof(T) == 4 ? N >= 2048 : N >= 4096) ? 16
                           : ((sizeof(T) == 8 && N == 1024) ? 4 : (N < 8 ? N : 8));
    static constexpr int TT = N / P;
    static constexpr int RMAX = 16;
};

}  // namespace shrt

// shrt_kern.hip -- kernel instances and launchers for one precision.
// Built twice: -DSHRT_REAL=float and -DSHRT_REAL=double.
#include <hip/hip_ext.h>

#include "launch.hpp"

#ifndef SHRT_REAL
#error "SHRT_REAL must be float or double"
#endif

namespace shrt {

namespace {

using R = SHRT_REAL;

// Transform lengths compiled in.  A kernel whose workgroups hold KT team
// buffers of length N fits when they fit the 160 KiB of LDS a CDNA4 workgroup
// can address: every length for one team (4096 in FP64 is 67.6 KB), all but
// 4096 in FP64 for the four-team tile kernels.
template <int N, int KT>
constexpr bool fits() {
    return (long long)KT * lds_len<N>() * (long long)sizeof(cplx<R>) <= 160 * 1024;
}

template <int KT = 1, typename F>
bool dispatch(int n, F&& f) {
    switch (n) {
#define SHRT_CASE(N)                                              \
    case N:                                                       \
        if constexpr (fits<N, KT>()) {                            \
            f(std::integral_constant<int, N>{});                  \
            return true;                                          \
        } else {                                                  \
            return false;                                         \
        }
        SHRT_CASE(32)
        SHRT_CASE(64)
        SHRT_CASE(128)
        SHRT_CASE(256)
        SHRT_CASE(512)
        SHRT_CASE(1024)
        SHRT_CASE(2048)
        SHRT_CASE(4096)
#undef SHRT_CASE
        default:
            return false;
    }
}

hipError_t unsupported() { return hipErrorInvalidValue; }

// A step-kernel launch; with timing events set (shrt_kernel_times) the
// dispatch itself records them at its start and end (hipExtLaunchKernelGGL),
// so the measured duration excludes the launch gaps around it.
template <typename F, typename... Args>
void step_launch(F kernel, const dim3& grid, const dim3& block, uint32_t shmem, hipStream_t s, Args... args) {
    if (g_launch_timing.start != nullptr)
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, g_launch_timing.start, g_launch_timing.stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
}

}  // namespace

template <>
bool Launch<R>::supported(int n) {
    return dispatch(n, [](auto) {});
}

template <>
bool Launch<R>::tiles_supported(int n) {
    return dispatch<kTile>(n, [](auto) {});
}

template <>
int Launch<R>::p2_workgroups(int nx, int ny) {
    int w = 0;
    dispatch(nx, [&](auto nc) { w = ny / p2_teams<R, decltype(nc)::value>(); });
    return w;
}

template <>
int Launch<R>::team_threads(int n) {
    int tt = 0;
    dispatch(n, [&](auto nc) { tt = TeamShape<R, decltype(nc)::value>::TT; });
    return tt;
}

template <>
hipError_t Launch<R>::p1(const StepArgs<R>& a, const StepArgs<R>* dargs, int count, int variant, hipStream_t s) {
    if (variant == kP1Tile) {
        const bool ok = dispatch<kTile>(a.g.ny, [&](auto nc) {
            constexpr int N = decltype(nc)::value;
            const P1Grid pg(a.g.nkxh, a.ph.linear);
            const dim3 grid(pg.grid, dargs ? count : 1), block(kTile * TeamShape<R, N>::TT);
            if (dargs) step_launch((k_p1<R, N, true>), grid, block, 0, s, a, dargs);
            else step_launch((k_p1<R, N, false>), grid, block, 0, s, a, dargs);
        });
        return ok ? hipGetLastError() : unsupported();
    }
    bool ok = dispatch(a.g.ny, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        if (variant == kP1Pair) {
            const dim3 grid(a.ph.linear ? a.g.nkxh : 3 * (a.g.nkxh - 1) + 2, dargs ? count : 1);
            if (dargs) step_launch((k_p1_pair<R, N, true>), grid, dim3(TeamShape<R, N>::TT), 0, s, a, dargs);
            else step_launch((k_p1_pair<R, N, false>), grid, dim3(TeamShape<R, N>::TT), 0, s, a, dargs);
        } else if (variant == kP1Row) {
            const dim3 grid((a.g.nkxh - 1) * (a.ph.linear ? 2 : 6) + (a.ph.linear ? 1 : 3), dargs ? count : 1);
            if (dargs) step_launch((k_p1_row<R, N, true>), grid, dim3(TeamShape<R, N>::TT), 0, s, a, dargs);
            else step_launch((k_p1_row<R, N, false>), grid, dim3(TeamShape<R, N>::TT), 0, s, a, dargs);
        }
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::p2(const StepArgs<R>& a, const StepArgs<R>* dargs, int count, int variant, hipStream_t s) {
    bool ok = dispatch(a.g.nx, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        constexpr int KT = p2_teams<R, N>();
        const dim3 grid(a.g.ny / KT, dargs ? count : 1), block(KT * Team<N, p2_points<R, N>()>::TT);
        if (variant == kP2Three) {
            if (dargs) step_launch((k_p2<R, N, true, true>), grid, block, 0, s, a, dargs);
            else step_launch((k_p2<R, N, false, true>), grid, block, 0, s, a, dargs);
        } else {
            if (dargs) step_launch((k_p2<R, N, true, false>), grid, block, 0, s, a, dargs);
            else step_launch((k_p2<R, N, false, false>), grid, block, 0, s, a, dargs);
        }
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::p3(const StepArgs<R>& a, const StepArgs<R>* dargs, int count, int variant, hipStream_t s) {
    auto go = [&](auto ktc) {
        constexpr int KT = decltype(ktc)::value;
        return dispatch<KT>(a.g.ny, [&](auto nc) {
            constexpr int N = decltype(nc)::value;
            const dim3 grid(2 * ((a.g.nkxh + KT - 1) / KT), dargs ? count : 1), block(KT * TeamShape<R, N>::TT);
            if (dargs) step_launch((k_p3<R, N, true, KT, false>), grid, block, 0, s, a, dargs);
            else step_launch((k_p3<R, N, false, KT, false>), grid, block, 0, s, a, dargs);
        });
    };
    if (sizeof(R) == 4 && a.ph.narrow) {   // the FP32-arithmetic update: one row pair per workgroup
        const bool ok = dispatch(a.g.ny, [&](auto nc) {
            constexpr int N = decltype(nc)::value;
            const dim3 grid(2 * a.g.nkxh, dargs ? count : 1), block(TeamShape<R, N>::TT);
            if (dargs) step_launch((k_p3<R, N, true, 1, false, true>), grid, block, 0, s, a, dargs);
            else step_launch((k_p3<R, N, false, 1, false, true>), grid, block, 0, s, a, dargs);
        });
        return ok ? hipGetLastError() : unsupported();
    }
    if (variant == kP3Split) {   // one row of a row pair per workgroup
        const bool ok = dispatch(a.g.ny, [&](auto nc) {
            constexpr int N = decltype(nc)::value;
            const dim3 grid(4 * a.g.nkxh, dargs ? count : 1), block(TeamShape<R, N>::TT);
            if (dargs) step_launch((k_p3<R, N, true, 1, false, false, true>), grid, block, 0, s, a, dargs);
            else step_launch((k_p3<R, N, false, 1, false, false, true>), grid, block, 0, s, a, dargs);
        });
        return ok ? hipGetLastError() : unsupported();
    }
    const bool ok = variant == kP3Pair ? go(std::integral_constant<int, 1>{}) : go(std::integral_constant<int, kTile>{});
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::p31(const StepArgs<R>& a, const StepArgs<R>* dargs, int count, hipStream_t s) {
    if (a.ph.narrow) return unsupported();   // the FP32-arithmetic update runs unfused
    bool ok = dispatch(a.g.ny, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        const dim3 grid(3 * (a.g.nkxh - 1) + 2, dargs ? count : 1), block(TeamShape<R, N>::TT);
        if (dargs) step_launch((k_p3<R, N, true, 1, true>), grid, block, 0, s, a, dargs);
        else step_launch((k_p3<R, N, false, 1, true>), grid, block, 0, s, a, dargs);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::xfwd(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.nx, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_xfwd<R, N>), dim3(a.g.ntile_y), dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::yfwd(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.ny, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_yfwd<R, N>), dim3(a.g.ntile_x), dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::yinv(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.ny, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_yinv<R, N>), dim3(a.g.ntile_x), dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::xinv(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.nx, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_xinv<R, N>), dim3(a.g.ntile_y), dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::yinv_lab(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.ny, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_yinv_lab<R, N>), dim3((a.g.nkx + kTile - 1) / kTile),
                           dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::xinv_lab(const AuxArgs<R>& a, hipStream_t s) {
    bool ok = dispatch<kTile>(a.g.nx, [&](auto nc) {
        constexpr int N = decltype(nc)::value;
        hipLaunchKernelGGL((k_aux_xinv_lab<R, N>), dim3(a.g.ntile_y), dim3(kTile * TeamShape<R, N>::TT), 0, s, a);
    });
    return ok ? hipGetLastError() : unsupported();
}

template <>
hipError_t Launch<R>::view_copy(const Geo& g, const void* in, void* out, const int* jmp, hipStream_t s) {
    const long long M = (long long)g.nkx * g.nky;
    hipLaunchKernelGGL((k_view_copy<R>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, g,
                       static_cast<const cplx<R>*>(in), static_cast<cplx<R>*>(out), jmp);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::phi_from_omega(const Geo& g, const void* om, void* phi, const double* kx, const double* ky,
                                     const double* sh, hipStream_t s) {
    const long long M = (long long)g.nkx * g.nky;
    hipLaunchKernelGGL((k_phi_from_omega<R>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, g,
                       static_cast<const cplx<R>*>(om), static_cast<cplx<R>*>(phi), kx, ky, sh);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::lowpass(const Geo& g, void* spec, int kcut, hipStream_t s) {
    const long long M = (long long)g.nkx * g.nky;
    hipLaunchKernelGGL((k_lowpass<R>), dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, g, static_cast<cplx<R>*>(spec),
                       kcut);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::fill_zero(void* p, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill_zero<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       static_cast<cplx<R>*>(p), n);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::mul_acc(const void* a, const void* b, void* out, long long n, int negate, int accumulate,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_mul_acc<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       static_cast<const R*>(a), static_cast<const R*>(b), static_cast<R*>(out), n, negate,
                       accumulate);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::net_in(const void* f, float* x, long long n, double scale, double fill, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_net_in<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, static_cast<const R*>(f), x,
                       n, scale, fill);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::net_out(const float* o, void* f, long long n, double scale, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_net_out<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, o, static_cast<R*>(f), n,
                       scale);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::sumsq(const void* a, const float* o, long long n, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_sumsq<R>), dim3(1), dim3(1024), 0, s, static_cast<const cplx<R>*>(a), o, n, out);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::axpy_round(const void* a, const void* c, double dt, void* dst, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_axpy_round<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       static_cast<const cplx<R>*>(a), static_cast<const cplx<R>*>(c), dt, static_cast<cplx<R>*>(dst),
                       n);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::copy_stream(const void* src, void* dst, long long nwords, hipStream_t s) {
    if (nwords <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_copy_stream<R>), dim3((unsigned)((nwords + 1023) / 1024)), dim3(256), 0, s,
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), nwords);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::copy16(const void* src, void* dst, long long nwords, hipStream_t s) {
    hipLaunchKernelGGL((k_copy16<R>), dim3(4096), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), nwords);
    return hipGetLastError();
}

template <>
hipError_t Launch<R>::diag(const Geo& g, const void* om, const void* rh, const int* jmp, const double* kx,
                           const double* ky, const double* sh, double* partial, int nblocks, double* out,
                           hipStream_t s) {
    hipLaunchKernelGGL((k_diag_partial<R>), dim3(nblocks), dim3(kDiagThreads), 0, s, g,
                       static_cast<const cplx<R>*>(om), static_cast<const cplx<R>*>(rh), jmp, kx, ky, sh, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_diag_final<R>), dim3(1), dim3(64), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

}  // namespace shrt

// launch.hpp -- host-side launchers of the device kernels, one explicit
// instantiation per precision (shrt_kern.hip is compiled once for float and
// once for double).  Each launcher picks the kernel instance for the transform
// length and returns the launch status.
#pragma once

#include "aux_kernels.hpp"
#include "step_kernels.hpp"

namespace shrt {

// Events the step-kernel launchers hand to the dispatch when set (per host
// thread; defined in shrt_api.hip).
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

template <typename T>
struct Launch {
    static bool supported(int n);         // the default (one-team) step kernels fit length n
    static bool tiles_supported(int n);   // the four-team kernels (P1/P3 tiles, the aux transforms) fit
    static int team_threads(int n);
    static int p2_workgroups(int nx, int ny);   // the P2 grid (number of CFL partials)
    // The step kernels.  dargs == nullptr: one instance, arguments `geom` passed
    // by value.  Otherwise `count` instances of one shape whose argument blocks
    // are in device memory at `dargs` (instance i runs as blockIdx.y == i);
    // `geom` is any one of them (host copy, for the grid).
    // `variant` is that pass's member of Variants.
    static hipError_t p1(const StepArgs<T>& geom, const StepArgs<T>* dargs, int count, int variant, hipStream_t s);
    static hipError_t p2(const StepArgs<T>& geom, const StepArgs<T>* dargs, int count, int variant, hipStream_t s);
    static hipError_t p3(const StepArgs<T>& geom, const StepArgs<T>* dargs, int count, int variant, hipStream_t s);
    // P3 of step n fused with P1 of step n+1 (one row pair per workgroup).
    static hipError_t p31(const StepArgs<T>& geom, const StepArgs<T>* dargs, int count, hipStream_t s);
    static hipError_t xfwd(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t yfwd(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t yinv(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t xinv(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t yinv_lab(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t xinv_lab(const AuxArgs<T>& a, hipStream_t s);
    static hipError_t view_copy(const Geo& g, const void* in, void* out, const int* jmp, hipStream_t s);
    static hipError_t phi_from_omega(const Geo& g, const void* om, void* phi, const double* kx,
                                     const double* ky, const double* sh, hipStream_t s);
    static hipError_t fill_zero(void* p, long long n, hipStream_t s);
    static hipError_t lowpass(const Geo& g, void* spec, int kcut, hipStream_t s);
    static hipError_t mul_acc(const void* a, const void* b, void* out, long long n, int negate, int accumulate,
                              hipStream_t s);
    static hipError_t copy16(const void* src, void* dst, long long nwords, hipStream_t s);
    static hipError_t copy_stream(const void* src, void* dst, long long nwords, hipStream_t s);
    // closure correction helpers (shrt_closure_correct)
    static hipError_t net_in(const void* f, float* x, long long n, double scale, double fill, hipStream_t s);
    static hipError_t net_out(const float* o, void* f, long long n, double scale, hipStream_t s);
    static hipError_t sumsq(const void* a, const float* o, long long n, double* out, hipStream_t s);
    static hipError_t axpy_round(const void* a, const void* c, double dt, void* dst, long long n, hipStream_t s);
    static hipError_t diag(const Geo& g, const void* om, const void* rh, const int* jmp, const double* kx,
                           const double* ky, const double* sh, double* partial, int nblocks, double* out,
                           hipStream_t s);
};

}  // namespace shrt

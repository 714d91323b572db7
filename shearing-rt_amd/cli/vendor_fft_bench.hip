// vendor_fft_bench.hip -- comparison only, not part of the library: the FFT
// work of one reference step done by the vendor library.  The reference's
// time_advance makes six 2-D complex-to-real transforms (the fields and
// derivatives of ktox, fft.cu:170-204) and two real-to-complex ones (the
// brackets of xtok, fft.cu:147-166) of nx x ny per step; this times the same
// transforms with hipFFT (rocFFT) plans on one stream, out of place, and
// prints one JSON line.  The library's own step does all of its transforms in
// its three kernels; bench.py --vendor-fft runs this for the comparison.
//
//   vendor_fft_bench [nx] [ny] [reps] [fp64]
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHECK_HIP(x)                                                                 \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
            return 1;                                                                \
        }                                                                            \
    } while (0)
#define CHECK_FFT(x)                                                                 \
    do {                                                                             \
        hipfftResult r_ = (x);                                                       \
        if (r_ != HIPFFT_SUCCESS) {                                                  \
            std::fprintf(stderr, "%s: hipfft error %d\n", #x, (int)r_);              \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char** argv) {
    const int nx = argc > 1 ? std::atoi(argv[1]) : 2048;
    const int ny = argc > 2 ? std::atoi(argv[2]) : 2048;
    const int reps = argc > 3 ? std::atoi(argv[3]) : 100;
    const bool dbl = argc > 4 && std::strcmp(argv[4], "fp64") == 0;
    if (nx < 2 || ny < 2 || reps < 1) {
        std::fprintf(stderr, "usage: vendor_fft_bench [nx] [ny] [reps] [fp64]\n");
        return 2;
    }
    const size_t real_elems = (size_t)nx * ny;
    const size_t cplx_elems = (size_t)nx * (ny / 2 + 1);
    const size_t rb = dbl ? sizeof(double) : sizeof(float);
    void *re = nullptr, *cx = nullptr;
    CHECK_HIP(hipMalloc(&re, real_elems * rb));
    CHECK_HIP(hipMalloc(&cx, cplx_elems * 2 * rb));
    CHECK_HIP(hipMemset(re, 0, real_elems * rb));
    CHECK_HIP(hipMemset(cx, 0, cplx_elems * 2 * rb));
    hipStream_t s;
    CHECK_HIP(hipStreamCreate(&s));
    hipfftHandle inv, fwd;
    CHECK_FFT(hipfftPlan2d(&inv, nx, ny, dbl ? HIPFFT_Z2D : HIPFFT_C2R));
    CHECK_FFT(hipfftPlan2d(&fwd, nx, ny, dbl ? HIPFFT_D2Z : HIPFFT_R2C));
    CHECK_FFT(hipfftSetStream(inv, s));
    CHECK_FFT(hipfftSetStream(fwd, s));
    auto c2r = [&]() {
        return dbl ? hipfftExecZ2D(inv, (hipfftDoubleComplex*)cx, (hipfftDoubleReal*)re)
                   : hipfftExecC2R(inv, (hipfftComplex*)cx, (hipfftReal*)re);
    };
    auto r2c = [&]() {
        return dbl ? hipfftExecD2Z(fwd, (hipfftDoubleReal*)re, (hipfftDoubleComplex*)cx)
                   : hipfftExecR2C(fwd, (hipfftReal*)re, (hipfftComplex*)cx);
    };
    for (int i = 0; i < 3; ++i) {   // warm-up
        CHECK_FFT(c2r());
        CHECK_FFT(r2c());
    }
    hipEvent_t e0, e1, e2;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipEventCreate(&e2));
    CHECK_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) CHECK_FFT(c2r());
    CHECK_HIP(hipEventRecord(e1, s));
    for (int i = 0; i < reps; ++i) CHECK_FFT(r2c());
    CHECK_HIP(hipEventRecord(e2, s));
    CHECK_HIP(hipEventSynchronize(e2));
    float t_inv = 0.f, t_fwd = 0.f;
    CHECK_HIP(hipEventElapsedTime(&t_inv, e0, e1));
    CHECK_HIP(hipEventElapsedTime(&t_fwd, e1, e2));
    const double c = t_inv / reps, f = t_fwd / reps;
    std::printf("{\"nx\": %d, \"ny\": %d, \"precision\": \"%s\", \"reps\": %d, \"ms_c2r\": %.6f, \"ms_r2c\": %.6f, "
                "\"ms_reference_step_ffts\": %.6f}\n",
                nx, ny, dbl ? "fp64" : "fp32", reps, c, f, 6.0 * c + 2.0 * f);
    hipfftDestroy(inv);
    hipfftDestroy(fwd);
    CHECK_HIP(hipFree(re));
    CHECK_HIP(hipFree(cx));
    return 0;
}

// shrt_api.hip -- host implementation of the C ABI in include/shrt.h.
//
// One context = one instance of the reference's global state (fields.cu:13-25,
// fft.cu:7-8, shear.cu:21-34), held in HBM and advanced by the three step
// kernels on the context's own stream.  K steps (params.graph_steps) are
// captured once into a HIP graph per starting parity and replayed; the device
// decides the CFL halvings, the shift advance and the remap itself, so no step
// needs the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/shrt.h"
#include "launch.hpp"

using namespace shrt;

namespace {

thread_local std::string g_last_error;

shrt_status fail(shrt_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(SHRT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kHalvingCap = 256;
constexpr int kDiagBlocks = 512;

}  // namespace

// Error text for entry points defined in other translation units (the closure).
namespace shrt {
void set_last_error(const std::string& msg) { g_last_error = msg; }
thread_local LaunchTiming g_launch_timing;
// closure_api.hip: device access to the network for shrt_closure_correct
int closure_grid(const shrt_closure* n);
float* closure_input(shrt_closure* n);
const float* closure_output(const shrt_closure* n);
hipError_t closure_forward_on(shrt_closure* n, hipStream_t s);
}  // namespace shrt

struct shrt_ctx {
    shrt_params p{};
    int prec = 4;
    size_t esz = 8;     // bytes per complex element
    size_t rsz = 4;     // bytes per real element
    Geo g{};
    Phys ph{};
    hipStream_t stream = nullptr;
    int device = 0;          // the device current at creation (the stream's)
    void* ring[4][3] = {};   // a_omg, d_omg, a_rho, d_rho
    void* aphi = nullptr;
    void* G = nullptr;
    void* H = nullptr;
    void* tmp_spec = nullptr;
    void* tmp_spec2 = nullptr;
    void* tmp_phys = nullptr;
    double* kx_d = nullptr;
    double* ky_d = nullptr;
    void* twx = nullptr;
    void* twy = nullptr;
    StepCtl* ctl = nullptr;   // [2]
    double* s = nullptr;      // [2][nkx]
    int* jump = nullptr;      // [2][nkx]
    int* jump_tmp = nullptr;  // [nkx]
    HalvingRec* hlog = nullptr;
    double* cflp = nullptr;   // [P2 workgroups][2] CFL partial maxima of the x pass
    Variants var;             // kernel variant of each pass (shrt_autotune)
    double* diag_partial = nullptr;
    double* diag_out = nullptr;
    int cur = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<double> kx_h, ky_h;
    std::vector<void*> allocs;
};

namespace {

// ---------------------------------------------------------------------------
// allocation and tables
// ---------------------------------------------------------------------------
shrt_status dalloc(shrt_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(SHRT_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    return SHRT_OK;
}

template <typename T>
std::vector<cplx<T>> twiddles(int n) {
    std::vector<cplx<T>> w(n);
    for (int k = 0; k < n; ++k) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        w[k] = {(T)cosl(ang), (T)sinl(ang)};
    }
    return w;
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

bool prec_supported(int prec, int n) {
    return prec == 4 ? Launch<float>::supported(n) : Launch<double>::supported(n);
}

// ---------------------------------------------------------------------------
// argument blocks
// ---------------------------------------------------------------------------
template <typename T>
StepArgs<T> step_args(shrt_ctx* c, int in_par, int store_phi, int store_phi_next = 0) {
    StepArgs<T> a{};
    a.g = c->g;
    a.ph = c->ph;
    for (int s = 0; s < 3; ++s) {
        a.a_omg[s] = static_cast<cplx<T>*>(c->ring[0][s]);
        a.d_omg[s] = static_cast<cplx<T>*>(c->ring[1][s]);
        a.a_rho[s] = static_cast<cplx<T>*>(c->ring[2][s]);
        a.d_rho[s] = static_cast<cplx<T>*>(c->ring[3][s]);
    }
    a.aphi = static_cast<cplx<T>*>(c->aphi);
    a.G = static_cast<cplx<T>*>(c->G);
    a.H = static_cast<cplx<T>*>(c->H);
    a.kx = c->kx_d;
    a.ky = c->ky_d;
    a.twx = static_cast<const cplx<T>*>(c->twx);
    a.twy = static_cast<const cplx<T>*>(c->twy);
    a.ctl_in = c->ctl + in_par;
    a.ctl_out = c->ctl + (in_par ^ 1);
    a.s_in = c->s + (size_t)in_par * c->g.nkx;
    a.s_out = c->s + (size_t)(in_par ^ 1) * c->g.nkx;
    a.jump = c->jump;
    a.hlog = c->hlog;
    a.hlog_cap = kHalvingCap;
    a.store_phi = store_phi;
    a.store_phi_next = store_phi_next;
    a.cflp = c->cflp;
    a.ncflp = Launch<T>::p2_workgroups(c->g.nx, c->g.ny);
    return a;
}

template <typename T>
AuxArgs<T> aux_args(shrt_ctx* c, const void* in, void* out, const double* s, const int* jmp, bool xdir, int mode) {
    AuxArgs<T> a{};
    a.g = c->g;
    a.in = in;
    a.out = out;
    a.kx = c->kx_d;
    a.ky = c->ky_d;
    a.s = s;
    a.jmp = jmp;
    a.tw = static_cast<const cplx<T>*>(xdir ? c->twx : c->twy);
    a.mode = mode;
    return a;
}

template <typename T>
hipError_t enqueue_step_t(shrt_ctx* c, int in_par, int store_phi) {
    const StepArgs<T> a = step_args<T>(c, in_par, store_phi);
    hipError_t e = Launch<T>::p1(a, nullptr, 1, c->var.p1, c->stream);
    if (e != hipSuccess) return e;
    e = Launch<T>::p2(a, nullptr, 1, c->var.p2, c->stream);
    if (e != hipSuccess) return e;
    return Launch<T>::p3(a, nullptr, 1, c->var.p3, c->stream);
}

hipError_t enqueue_step(shrt_ctx* c, int in_par, int store_phi) {
    return c->prec == 4 ? enqueue_step_t<float>(c, in_par, store_phi) : enqueue_step_t<double>(c, in_par, store_phi);
}

// Step k of a captured batch of K >= 2 steps.  Fused: step 0 runs P1, P2 and
// P3+P1 of step 1; steps 1..K-2 run P2 and P3+P1 of the next step; step K-1
// runs P2 and P3.  aphi is stored for the batch's last step (by the fused
// kernel of step K-2), as the unfused batch stores it in that step's P1.
template <typename T>
hipError_t enqueue_batch_step_t(shrt_ctx* c, int in_par, int k, int K, hipStream_t s) {
    const bool last = k == K - 1;
    if (!c->var.fuse || c->ph.narrow) {
        const StepArgs<T> a = step_args<T>(c, in_par, last);
        hipError_t e = Launch<T>::p1(a, nullptr, 1, c->var.p1, s);
        if (e == hipSuccess) e = Launch<T>::p2(a, nullptr, 1, c->var.p2, s);
        if (e == hipSuccess) e = Launch<T>::p3(a, nullptr, 1, c->var.p3, s);
        return e;
    }
    const int sp = k + 1 == K - 1;
    const StepArgs<T> a = step_args<T>(c, in_par, 0, sp);
    hipError_t e = k == 0 ? Launch<T>::p1(a, nullptr, 1, c->var.p1, s) : hipSuccess;
    if (e == hipSuccess) e = Launch<T>::p2(a, nullptr, 1, c->var.p2, s);
    if (e == hipSuccess) e = last ? Launch<T>::p3(a, nullptr, 1, c->var.p3, s) : Launch<T>::p31(a, nullptr, 1, s);
    return e;
}

shrt_status build_graph(shrt_ctx* c, int par) {
    if (c->gexec[par]) return SHRT_OK;
    const int K = c->p.graph_steps;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int k = 0; k < K && err == hipSuccess; ++k)
        err = c->prec == 4 ? enqueue_batch_step_t<float>(c, par ^ (k & 1), k, K, c->stream)
                           : enqueue_batch_step_t<double>(c, par ^ (k & 1), k, K, c->stream);
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (err != hipSuccess) return fail(SHRT_ERR_DEVICE, std::string("capture: ") + hipGetErrorString(err));
    if (e2 != hipSuccess) return fail(SHRT_ERR_DEVICE, std::string("end capture: ") + hipGetErrorString(e2));
    hipError_t e3 = hipGraphInstantiate(&c->gexec[par], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) {
        c->gexec[par] = nullptr;
        return fail(SHRT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e3));
    }
    return SHRT_OK;
}

// ---------------------------------------------------------------------------
// control-block helpers (host side, synchronous)
// ---------------------------------------------------------------------------
shrt_status read_ctl(shrt_ctx* c, StepCtl* out) {
    HIP_TRY(hipMemcpyAsync(out, c->ctl + c->cur, sizeof(StepCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status write_ctl(shrt_ctx* c, const StepCtl& v) {
    HIP_TRY(hipMemcpyAsync(c->ctl + c->cur, &v, sizeof(StepCtl), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

// Remap view of a buffer whose frame is `buf_epoch` at control state `st`.
shrt_status view_for(shrt_ctx* c, const StepCtl& st, int buf_epoch, const int** jmp) {
    *jmp = nullptr;
    if (buf_epoch == st.epoch) return SHRT_OK;
    if (st.epoch - buf_epoch != 1)
        return fail(SHRT_ERR_STATE, "buffer is more than one remap behind the current frame");
    *jmp = c->jump + (size_t)(st.epoch & 1) * c->g.nkx;
    return SHRT_OK;
}

hipError_t view_copy(shrt_ctx* c, const void* in, void* out, const int* jmp) {
    return c->prec == 4 ? Launch<float>::view_copy(c->g, in, out, jmp, c->stream)
                        : Launch<double>::view_copy(c->g, in, out, jmp, c->stream);
}

hipError_t phi_from_omega(shrt_ctx* c, const void* om, void* phi, const double* s) {
    return c->prec == 4 ? Launch<float>::phi_from_omega(c->g, om, phi, c->kx_d, c->ky_d, s, c->stream)
                        : Launch<double>::phi_from_omega(c->g, om, phi, c->kx_d, c->ky_d, s, c->stream);
}

size_t spec_bytes(const shrt_ctx* c) { return (size_t)c->g.nkx * c->g.nky * c->esz; }
size_t phys_bytes(const shrt_ctx* c) { return (size_t)c->g.nx * c->g.ny * c->rsz; }

// The omega and rho buffers of an a-ring slot share one frame record, and so
// do the two d-ring buffers of a slot.  Before one buffer of a slot is
// overwritten in the current frame, its sibling is brought into that frame
// too (the remap view applied in place through the scratch buffer), so that
// marking the slot current does not drop the sibling's pending remap.
shrt_status materialise(shrt_ctx* c, const StepCtl& v, int ring, int slot) {
    const int ep = (ring == 0 || ring == 2) ? v.slot_epoch_a[slot] : v.slot_epoch_d[slot];
    const int* jmp;
    shrt_status st = view_for(c, v, ep, &jmp);
    if (st != SHRT_OK || jmp == nullptr) return st;
    HIP_TRY(view_copy(c, c->ring[ring][slot], c->tmp_spec, jmp));
    HIP_TRY(hipMemcpyAsync(c->ring[ring][slot], c->tmp_spec, spec_bytes(c), hipMemcpyDeviceToDevice, c->stream));
    return SHRT_OK;
}

// Physical slot of the reference's logical ring name (a0/a1/a2 after the last
// rotation) when `n` steps are complete.
int logical_slot(long long n, int j) {
    return j == 0 ? slot_a0(n - 1) : j == 1 ? slot_a1(n - 1) : slot_a2(n - 1);
}

// xtok on device buffers (phys -> spec), via H field 0 scratch.
template <typename T>
hipError_t xtok_dev(shrt_ctx* c, const void* phys, void* spec) {
    AuxArgs<T> a = aux_args<T>(c, phys, c->H, nullptr, nullptr, true, 0);
    hipError_t e = Launch<T>::xfwd(a, c->stream);
    if (e != hipSuccess) return e;
    AuxArgs<T> b = aux_args<T>(c, c->H, spec, nullptr, nullptr, false, 0);
    return Launch<T>::yfwd(b, c->stream);
}
hipError_t xtok_dev(shrt_ctx* c, const void* phys, void* spec) {
    return c->prec == 4 ? xtok_dev<float>(c, phys, spec) : xtok_dev<double>(c, phys, spec);
}

// ktox on device buffers (spec -> phys) with an optional multiplier and view.
template <typename T>
hipError_t ktox_dev(shrt_ctx* c, const void* spec, void* phys, const double* s, const int* jmp, int mode) {
    AuxArgs<T> a = aux_args<T>(c, spec, c->G, s, jmp, false, mode);
    hipError_t e = Launch<T>::yinv(a, c->stream);
    if (e != hipSuccess) return e;
    AuxArgs<T> b = aux_args<T>(c, c->G, phys, s, nullptr, true, 0);
    return Launch<T>::xinv(b, c->stream);
}
hipError_t ktox_dev(shrt_ctx* c, const void* spec, void* phys, const double* s, const int* jmp, int mode) {
    return c->prec == 4 ? ktox_dev<float>(c, spec, phys, s, jmp, mode) : ktox_dev<double>(c, spec, phys, s, jmp, mode);
}

// Lab-frame ktox_shear on device buffers (spec -> phys) with the shift `s` and
// an optional remap view; the nkx x ny intermediate lives in G.
template <typename T>
hipError_t ktox_shear_dev(shrt_ctx* c, const void* spec, void* phys, const double* s, const int* jmp) {
    AuxArgs<T> a = aux_args<T>(c, spec, c->G, s, jmp, false, 0);
    hipError_t e = Launch<T>::yinv_lab(a, c->stream);
    if (e != hipSuccess) return e;
    AuxArgs<T> b = aux_args<T>(c, c->G, phys, s, nullptr, true, 0);
    return Launch<T>::xinv_lab(b, c->stream);
}
hipError_t ktox_shear_dev(shrt_ctx* c, const void* spec, void* phys, const double* s, const int* jmp) {
    return c->prec == 4 ? ktox_shear_dev<float>(c, spec, phys, s, jmp) : ktox_shear_dev<double>(c, spec, phys, s, jmp);
}

hipError_t mul_acc(shrt_ctx* c, const void* a, const void* b, void* out, long long n, int negate, int accumulate) {
    return c->prec == 4 ? Launch<float>::mul_acc(a, b, out, n, negate, accumulate, c->stream)
                        : Launch<double>::mul_acc(a, b, out, n, negate, accumulate, c->stream);
}

hipError_t zero_buffer(shrt_ctx* c, void* p, long long nelem) {
    return c->prec == 4 ? Launch<float>::fill_zero(p, nelem, c->stream)
                        : Launch<double>::fill_zero(p, nelem, c->stream);
}

StepCtl fresh_ctl(const shrt_ctx* c) {
    StepCtl v{};
    v.time = 0.0;
    v.delt = c->p.delt;
    v.istep = 0;
    v.epoch = 0;
    v.last_remap_step = -1;
    return v;
}

// Reset every buffer and the control state (the zeroed allocation of
// fields.cu:120-166 and init_shear, shear.cu:149-185).
shrt_status reset_state(shrt_ctx* c) {
    const long long M = (long long)c->g.nkx * c->g.nky;
    for (int r = 0; r < 4; ++r)
        for (int s = 0; s < 3; ++s) HIP_TRY(zero_buffer(c, c->ring[r][s], M));
    HIP_TRY(zero_buffer(c, c->aphi, M));
    HIP_TRY(hipMemsetAsync(c->s, 0, sizeof(double) * 2 * c->g.nkx, c->stream));
    HIP_TRY(hipMemsetAsync(c->jump, 0, sizeof(int) * 2 * c->g.nkx, c->stream));
    StepCtl v = fresh_ctl(c);
    HIP_TRY(hipMemcpyAsync(c->ctl, &v, sizeof(StepCtl), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ctl + 1, &v, sizeof(StepCtl), hipMemcpyHostToDevice, c->stream));
    c->cur = 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The reference IC on the host in the context precision (time_integral.cu:236-248)
template <typename T>
void reference_ic_phys(const shrt_ctx* c, std::vector<T>& omgz, std::vector<T>& rho) {
    const int nx = c->g.nx, ny = c->g.ny;
    const T dx = (T)c->p.Lx / (T)nx, dy = (T)c->p.Ly / (T)ny;
    std::vector<T> xx(nx + 1), yy(ny + 1);
    for (int ix = 0; ix <= nx; ix++) xx[ix] = ix * dx;
    for (int iy = 0; iy <= ny; iy++) yy[iy] = iy * dy;
    omgz.assign((size_t)nx * ny, T(0));
    rho.assign((size_t)nx * ny, T(0));
    const T eps1 = (T)c->p.rho_eps1;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) rho[(size_t)ix * ny + iy] = eps1 * std::cos(2 * xx[ix] + 10 * yy[iy]);
}

// init_dis noise (time_integral.cu:261-307): glibc rand() after srand(2000),
// ikx-major, two draws per masked mode, added to rho_hat.
template <typename T>
void reference_noise(const shrt_ctx* c, std::vector<cplx<T>>& rh) {
    const int nkx = c->g.nkx, nky = c->g.nky;
    const T eps2 = (T)(0.1 * (T)c->p.rho_eps1);
    std::vector<T> kx(nkx), ky(nky);
    for (int i = 0; i < nkx; ++i) kx[i] = (T)c->kx_h[i];
    for (int i = 0; i < nky; ++i) ky[i] = (T)c->ky_h[i];
    srand(2000);
    for (int ikx = 0; ikx < nkx; ikx++)
        for (int iky = 0; iky < nky; iky++) {
            T f1 = 0, f2 = 0;
            if (std::fabs(kx[ikx]) <= 2 && std::fabs(ky[iky]) != 0 && std::fabs(ky[iky]) <= 10) {
                f1 = (T)(eps2 * (2 * ((double)rand() / (1.0 + RAND_MAX)) - 1.0));
                f2 = (T)(eps2 * (2 * ((double)rand() / (1.0 + RAND_MAX)) - 1.0));
            }
            cplx<T>& r = rh[(size_t)ikx * nky + iky];
            r.x = r.x + f1;
            r.y = r.y + f2;
        }
}

template <typename T>
shrt_status initialize_reference(shrt_ctx* c) {
    shrt_status st = reset_state(c);
    if (st != SHRT_OK) return st;
    std::vector<T> omgz, rho;
    reference_ic_phys<T>(c, omgz, rho);
    const int slot = slot_a1(0);
    HIP_TRY(hipMemcpyAsync(c->tmp_phys, omgz.data(), phys_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(xtok_dev(c, c->tmp_phys, c->ring[0][slot]));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpyAsync(c->tmp_phys, rho.data(), phys_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(xtok_dev(c, c->tmp_phys, c->ring[2][slot]));
    if (c->p.noise) {
        std::vector<cplx<T>> rh((size_t)c->g.nkx * c->g.nky);
        HIP_TRY(hipMemcpyAsync(rh.data(), c->ring[2][slot], spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        reference_noise<T>(c, rh);
        HIP_TRY(hipMemcpyAsync(c->ring[2][slot], rh.data(), spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(phi_from_omega(c, c->ring[0][slot], c->aphi, c->s));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

template <typename T>
shrt_status initialize_seeded(shrt_ctx* c, const shrt_seeded_ic* ic) {
    shrt_status st = reset_state(c);
    if (st != SHRT_OK) return st;
    const int nkx = c->g.nkx, nky = c->g.nky, nkxh = c->g.nkxh;
    std::vector<cplx<T>> rh((size_t)nkx * nky, cplx<T>{T(0), T(0)});
    for (int ikx = 0; ikx < nkx; ++ikx) {
        const double kxv = (ikx < nkxh ? ikx : ikx - nkx) * 2 * M_PI / c->p.Lx;
        for (int iky = 0; iky < nky; ++iky) {
            if (iky == 0 && ikx >= nkxh) continue;
            const double kyv = iky * 2 * M_PI / c->p.Ly;
            const double km = std::sqrt(kxv * kxv + kyv * kyv);
            if (km < ic->kmin || km > ic->kmax || km == 0.0) continue;
            const uint64_t z = splitmix64(ic->seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(ikx * nky + iky + 1)));
            const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            const double th = 2 * M_PI * u;
            const double a = ic->amp * std::pow(km, ic->expo);
            const double re = a * std::cos(th), im = a * std::sin(th);
            rh[(size_t)ikx * nky + iky] = {(T)re, (T)im};
            if (iky == 0 && ikx > 0) rh[(size_t)(nkx - ikx) * nky] = {(T)re, (T)(-im)};
        }
    }
    const int slot = slot_a1(0);
    HIP_TRY(hipMemcpyAsync(c->ring[2][slot], rh.data(), spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// Average milliseconds of one device-to-device copy of `bytes` between two
// fresh buffers: the plain grid-stride copy (the HBM-counter reference) or the
// streaming copy (the achievable copy rate).
static shrt_status copy_bench(long long bytes, int reps, bool stream, double* ms) {
    if (bytes <= 0 || bytes % 16 != 0 || reps <= 0 || !ms) return fail(SHRT_ERR_INVALID, "bad arguments");
    if (bytes / 16 > 1024LL * 0xffffffffLL) return fail(SHRT_ERR_INVALID, "copy too large");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    shrt_status rc = SHRT_OK;
    auto launch = [&]() {
        return stream ? Launch<float>::copy_stream(a, b, bytes / 16, st) : Launch<float>::copy16(a, b, bytes / 16, st);
    };
    float t = 0.f;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        rc = fail(SHRT_ERR_DEVICE, "hipStreamCreate failed");
    } else if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = fail(SHRT_ERR_ALLOC, "copy bench allocation failed");
    } else if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
               hipMemsetAsync(a, 1, bytes, st) != hipSuccess || launch() != hipSuccess ||
               hipEventRecord(e0, st) != hipSuccess) {
        rc = fail(SHRT_ERR_DEVICE, "copy bench set-up failed");
    } else {
        for (int i = 0; i < reps && rc == SHRT_OK; ++i)
            if (launch() != hipSuccess) rc = fail(SHRT_ERR_DEVICE, "copy kernel launch failed");
        if (rc == SHRT_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = fail(SHRT_ERR_DEVICE, "copy bench timing failed");
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (st) (void)hipStreamDestroy(st);
    if (rc == SHRT_OK) *ms = t / reps;
    return rc;
}

extern "C" {

const char* shrt_last_error(void) { return g_last_error.c_str(); }
const char* shrt_version(void) { return "shrt-mi355x 0.1"; }

int shrt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

shrt_status shrt_sizes(int nx, int ny, int* o) {
    if (!o || nx < 1 || ny < 1) return fail(SHRT_ERR_INVALID, "shrt_sizes: bad arguments");
    const int nkx = (nx - 1) / 3 * 2 + 1, nky = (ny - 1) / 3 + 1, nkxh = (nkx + 1) / 2;
    o[0] = nkx; o[1] = nky; o[2] = nkxh; o[3] = nx - nkxh; o[4] = nx - nkxh * 2 + 1; o[5] = ny / 2 + 1;
    o[6] = nkx * nky; o[7] = nkxh * ny;
    return SHRT_OK;
}

int shrt_supported(int nx, int ny, int precision) {
    if (precision != SHRT_FP32 && precision != SHRT_FP64) return 0;
    if (!is_pow2(nx) || !is_pow2(ny)) return 0;
    return prec_supported(precision, nx) && prec_supported(precision, ny) ? 1 : 0;
}

shrt_status shrt_create(const shrt_params* p, shrt_ctx** out) {
    if (!p || !out) return fail(SHRT_ERR_INVALID, "shrt_create: null argument");
    *out = nullptr;
    if (p->precision != SHRT_FP32 && p->precision != SHRT_FP64)
        return fail(SHRT_ERR_INVALID, "precision must be SHRT_FP32 (4) or SHRT_FP64 (8)");
    if (!shrt_supported(p->nx, p->ny, p->precision))
        return fail(SHRT_ERR_UNSUPPORTED, "grid " + std::to_string(p->nx) + "x" + std::to_string(p->ny) +
                                              " is not a supported power-of-two size for this precision");
    if (p->graph_steps < 0 || (p->graph_steps % 2) != 0)
        return fail(SHRT_ERR_INVALID, "graph_steps must be 0 or an even number");
    if (!(p->delt > 0) || !(p->Lx > 0) || !(p->Ly > 0) || p->rho0 == 0)
        return fail(SHRT_ERR_INVALID, "delt, Lx, Ly must be positive and rho0 nonzero");
    int sz[8];
    shrt_sizes(p->nx, p->ny, sz);
    if (p->sigma != 0 && sz[2] > sz[1])
        return fail(SHRT_ERR_INVALID, "remap needs nkxh <= nky (nx <= ny) when sigma != 0");
    if (std::fabs(p->sigma) * p->delt * 3.0 > 1.0)
        return fail(SHRT_ERR_INVALID, "|sigma|*delt must be below 1/3 (remaps would overlap)");

    shrt_ctx* c = new shrt_ctx();
    c->p = *p;
    c->prec = p->precision;
    c->rsz = c->prec;
    c->esz = 2 * c->prec;
    Geo& g = c->g;
    g.nx = p->nx; g.ny = p->ny;
    g.nkx = sz[0]; g.nky = sz[1]; g.nkxh = sz[2]; g.nkxh2 = sz[3]; g.nkxpad = sz[4]; g.ncy = sz[5];
    g.ntile_x = (g.nkxh + kTile - 1) / kTile;
    g.ntile_y = g.ny / kTile;
    g.I = (long long)g.nkxh * g.ny;
    g.Lx = p->Lx; g.Ly = p->Ly;
    g.dx = p->Lx / p->nx; g.dy = p->Ly / p->ny;
    g.scale = 1.0 / ((double)p->nx * (double)p->ny);
    Phys& ph = c->ph;
    ph.nu = p->nu; ph.kappa = p->kappa; ph.sigma = p->sigma; ph.g = p->g; ph.rho0 = p->rho0;
    ph.rho0_prime = p->rho0_prime; ph.cfl_num = p->cfl_num; ph.eps = 1e-10; ph.linear = p->linear ? 1 : 0;
    ph.narrow = 0;
    // Default P1 variant by the y length, as shrt_autotune picks it on MI355X:
    // one (row pair, field group) per workgroup at 2048 in FP32 and at 4096, the
    // four-row tile at 512, one (row pair, output) otherwise (the 8-point teams
    // of 1024 and of FP64 2048 fill the chip with one output per workgroup).
    if (g.ny >= 4096 || (g.ny == 2048 && c->prec == 4)) c->var.p1 = kP1Pair;
    else if (g.ny == 512) c->var.p1 = kP1Tile;
    // Captured batches fuse each step's P3 with the next step's P1 by default
    // where that measured faster on MI355X: FP32 at 512, 1024 and 2048 (512^2
    // 23.1 against 23.4 us per step, 1024^2 36.1 against 37.1-37.5, 2048^2
    // 77.0-77.2 against 80.8-81.9, one box); slower at 256, at 4096 and in FP64.
    if (c->prec == 4 && (g.ny == 512 || g.ny == 1024 || g.ny == 2048)) c->var.fuse = 1;

    auto bail = [&](shrt_status st) {
        shrt_destroy(c);
        return st;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(SHRT_ERR_DEVICE, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SHRT_ERR_DEVICE, "hipStreamCreate failed"));
    const size_t M = (size_t)g.nkx * g.nky;
    shrt_status st;
    for (int r = 0; r < 4; ++r)
        for (int s = 0; s < 3; ++s)
            if ((st = dalloc(c, &c->ring[r][s], M * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->aphi, M * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->G, 6 * (size_t)g.I * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->H, 2 * (size_t)g.I * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->tmp_spec, M * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->tmp_spec2, M * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->tmp_phys, (size_t)g.nx * g.ny * c->rsz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->kx_d, sizeof(double) * g.nkx)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->ky_d, sizeof(double) * g.nky)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->twx, (size_t)g.nx * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, &c->twy, (size_t)g.ny * c->esz)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->ctl, 2 * sizeof(StepCtl))) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->s, 2 * sizeof(double) * g.nkx)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->jump, 2 * sizeof(int) * g.nkx)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->jump_tmp, sizeof(int) * g.nkx)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->hlog, sizeof(HalvingRec) * kHalvingCap)) != SHRT_OK) return bail(st);
    if ((st = dalloc(c, (void**)&c->cflp, sizeof(double) * 2 * g.ny)) != SHRT_OK) return bail(st);   // <= one per column
    if ((st = dalloc(c, (void**)&c->diag_partial, sizeof(double) * 2 * kDiagQ * kDiagBlocks)) != SHRT_OK)
        return bail(st);
    if ((st = dalloc(c, (void**)&c->diag_out, sizeof(double) * 4)) != SHRT_OK) return bail(st);

    // k tables exactly as fourier.cu:123-128 builds them, kept in FP64.
    c->kx_h.assign(g.nkx, 0.0);
    for (int ikx = 0; ikx < g.nkxh; ikx++) c->kx_h[ikx] = ikx * 2 * M_PI / p->Lx;
    for (int ikx = 0; ikx < g.nkxh - 1; ikx++) c->kx_h[g.nkx - ikx - 1] = -c->kx_h[ikx + 1];
    c->ky_h.assign(g.nky, 0.0);
    for (int iky = 0; iky < g.nky; iky++) c->ky_h[iky] = iky * 2 * M_PI / p->Ly;
    auto up = [&](void* d, const void* h, size_t n) { return hipMemcpy(d, h, n, hipMemcpyHostToDevice); };
    bool ok = up(c->kx_d, c->kx_h.data(), sizeof(double) * g.nkx) == hipSuccess &&
              up(c->ky_d, c->ky_h.data(), sizeof(double) * g.nky) == hipSuccess;
    if (c->prec == 4) {
        auto wx = twiddles<float>(g.nx), wy = twiddles<float>(g.ny);
        ok = ok && up(c->twx, wx.data(), wx.size() * sizeof(wx[0])) == hipSuccess &&
             up(c->twy, wy.data(), wy.size() * sizeof(wy[0])) == hipSuccess;
    } else {
        auto wx = twiddles<double>(g.nx), wy = twiddles<double>(g.ny);
        ok = ok && up(c->twx, wx.data(), wx.size() * sizeof(wx[0])) == hipSuccess &&
             up(c->twy, wy.data(), wy.size() * sizeof(wy[0])) == hipSuccess;
    }
    if (!ok) return bail(fail(SHRT_ERR_DEVICE, "table upload failed"));
    if ((st = reset_state(c)) != SHRT_OK) return bail(st);
    *out = c;
    return SHRT_OK;
}

shrt_status shrt_destroy(shrt_ctx* c) {
    if (!c) return SHRT_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; ++i)
        if (c->gexec[i]) (void)hipGraphExecDestroy(c->gexec[i]);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SHRT_OK;
}

shrt_status shrt_initialize(shrt_ctx* c) {
    if (!c) return fail(SHRT_ERR_INVALID, "null context");
    return c->prec == 4 ? initialize_reference<float>(c) : initialize_reference<double>(c);
}

shrt_status shrt_initialize_seeded(shrt_ctx* c, const shrt_seeded_ic* ic) {
    if (!c || !ic) return fail(SHRT_ERR_INVALID, "null argument");
    return c->prec == 4 ? initialize_seeded<float>(c, ic) : initialize_seeded<double>(c, ic);
}

shrt_status shrt_step_async(shrt_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return fail(SHRT_ERR_INVALID, "shrt_step: bad arguments");
    const int K = c->p.graph_steps;
    while (nsteps > 0) {
        if (K >= 2 && nsteps >= K) {
            shrt_status st = build_graph(c, c->cur);
            if (st != SHRT_OK) return st;
            HIP_TRY(hipGraphLaunch(c->gexec[c->cur], c->stream));
            nsteps -= K;   // K is even: the parity is unchanged
        } else {
            HIP_TRY(enqueue_step(c, c->cur, nsteps == 1));
            c->cur ^= 1;
            nsteps -= 1;
        }
    }
    return SHRT_OK;
}

shrt_status shrt_synchronize(shrt_ctx* c) {
    if (!c) return fail(SHRT_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    if (v.error) return fail(SHRT_ERR_STATE, "device reported a remap while a history buffer was still pending");
    return SHRT_OK;
}

shrt_status shrt_step(shrt_ctx* c, int nsteps) {
    shrt_status st = shrt_step_async(c, nsteps);
    if (st != SHRT_OK) return st;
    return shrt_synchronize(c);
}

shrt_status shrt_time_steps(shrt_ctx* c, int nsteps, double* ms_total) {
    if (!c || !ms_total || nsteps < 0) return fail(SHRT_ERR_INVALID, "bad arguments");
    // build the graphs outside the timed region
    if (c->p.graph_steps >= 2 && nsteps >= c->p.graph_steps) {
        shrt_status st = build_graph(c, c->cur);
        if (st != SHRT_OK) return st;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    shrt_status st = shrt_step_async(c, nsteps);
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != SHRT_OK) return st;
    *ms_total = ms;
    return shrt_synchronize(c);
}

shrt_status shrt_copy_bench(long long bytes, int reps, double* ms) { return copy_bench(bytes, reps, false, ms); }

shrt_status shrt_stream_copy_bench(long long bytes, int reps, double* ms) { return copy_bench(bytes, reps, true, ms); }

}  // extern "C"

namespace {

// Average duration (ms) of `reps` back-to-back launches of one pass's variant on
// context `t`, from the step state it holds (the passes are idempotent for a
// fixed input state).
template <typename T>
shrt_status time_pass(shrt_ctx* t, int pass, int variant, int reps, double* ms) {
    const StepArgs<T> a = step_args<T>(t, t->cur, 0);
    auto launch = [&]() {
        return pass == 0 ? Launch<T>::p1(a, nullptr, 1, variant, t->stream)
             : pass == 1 ? Launch<T>::p2(a, nullptr, 1, variant, t->stream)
                         : Launch<T>::p3(a, nullptr, 1, variant, t->stream);
    };
    HIP_TRY(launch());   // warm
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, t->stream));
    hipError_t err = hipSuccess;
    for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch();
    HIP_TRY(hipEventRecord(e1, t->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(err);
    *ms = f / reps;
    return SHRT_OK;
}

}  // namespace

extern "C" {

shrt_status shrt_autotune(shrt_ctx* c, int reps, shrt_tuning* out) {
    if (!c || reps < 1) return fail(SHRT_ERR_INVALID, "shrt_autotune: bad arguments");
    // a scratch instance of the same shape and parameters, a few steps into a seeded run
    shrt_ctx* t = nullptr;
    shrt_params p = c->p;
    const int K = c->p.graph_steps >= 2 ? c->p.graph_steps : 20;
    p.graph_steps = K;
    shrt_status st = shrt_create(&p, &t);
    if (st != SHRT_OK) return st;
    shrt_seeded_ic ic{12345u, 2.0, std::max(2.0, c->g.ny / 16.0), 1e-4, -1.5};
    st = shrt_initialize_seeded(t, &ic);
    if (st == SHRT_OK) st = shrt_step(t, 4);
    static const int nvar[3] = {3, 2, 3};
    double ms[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const bool tiles = c->prec == 4 ? Launch<float>::tiles_supported(c->g.ny) : Launch<double>::tiles_supported(c->g.ny);
    for (int pass = 0; pass < 3 && st == SHRT_OK; ++pass)
        for (int v = 0; v < nvar[pass] && st == SHRT_OK; ++v) {
            if (!tiles && ((pass == 0 && v == kP1Tile) || (pass == 2 && v == kP3Tile))) {
                ms[pass][v] = 1e300;   // four-team tiles do not fit this length's LDS
                continue;
            }
            st = c->prec == 4 ? time_pass<float>(t, pass, v, reps, &ms[pass][v])
                              : time_pass<double>(t, pass, v, reps, &ms[pass][v]);
        }
    Variants chosen;
    chosen.p1 = kP1Row;
    for (int v = 1; v < nvar[0]; ++v)
        if (ms[0][v] < ms[0][chosen.p1]) chosen.p1 = v;
    chosen.p2 = ms[1][1] < ms[1][0] ? kP2Four : kP2Three;
    chosen.p3 = kP3Pair;
    for (int v = 1; v < nvar[2]; ++v)
        if (ms[2][v] < ms[2][chosen.p3]) chosen.p3 = v;
    // captured batches with and without the P3+P1 fusion, alternated, best of two each
    double step_ms[2] = {1e300, 1e300};
    const int nb = std::max(2, reps / 4);
    for (int round = 0; round < 2 && st == SHRT_OK; ++round)
        for (int f = 0; f < 2 && st == SHRT_OK; ++f) {
            t->var = chosen;
            st = shrt_set_fusion(t, f);
            double ms_total = 0.0;
            if (st == SHRT_OK) st = shrt_step(t, K);   // capture and warm
            if (st == SHRT_OK) st = shrt_time_steps(t, K * nb, &ms_total);
            if (st == SHRT_OK) step_ms[f] = std::min(step_ms[f], ms_total / (K * nb));
        }
    (void)shrt_destroy(t);
    if (st != SHRT_OK) return st;
    chosen.fuse = step_ms[1] < step_ms[0] ? 1 : 0;
    c->var = chosen;
    for (int i = 0; i < 2; ++i)   // captured graphs hold the old launches
        if (c->gexec[i]) {
            (void)hipGraphExecDestroy(c->gexec[i]);
            c->gexec[i] = nullptr;
        }
    if (out) {
        out->p1 = c->var.p1;
        out->p2 = c->var.p2;
        out->p3 = c->var.p3;
        int k = 0;
        for (int pass = 0; pass < 3; ++pass)
            for (int v = 0; v < nvar[pass]; ++v) out->ms[k++] = ms[pass][v];
        out->fuse = c->var.fuse;
        out->step_ms[0] = step_ms[0];
        out->step_ms[1] = step_ms[1];
    }
    return SHRT_OK;
}

shrt_status shrt_set_variants(shrt_ctx* c, int p1, int p2, int p3) {
    if (!c || p1 < 0 || p1 > 2 || p2 < 0 || p2 > 1 || p3 < 0 || p3 > 2) return fail(SHRT_ERR_INVALID, "bad variant");
    const bool tiles = c->prec == 4 ? Launch<float>::tiles_supported(c->g.ny) : Launch<double>::tiles_supported(c->g.ny);
    if (!tiles && (p1 == kP1Tile || p3 == kP3Tile))
        return fail(SHRT_ERR_UNSUPPORTED, "the four-team tile variants do not fit this grid's LDS (FP64 4096)");
    c->var.p1 = p1;
    c->var.p2 = p2;
    c->var.p3 = p3;
    for (int i = 0; i < 2; ++i)
        if (c->gexec[i]) {
            (void)hipGraphExecDestroy(c->gexec[i]);
            c->gexec[i] = nullptr;
        }
    return SHRT_OK;
}

shrt_status shrt_set_update_precision(shrt_ctx* c, int wide) {
    if (!c || (wide != 0 && wide != 1)) return fail(SHRT_ERR_INVALID, "shrt_set_update_precision: bad arguments");
    c->ph.narrow = (c->prec == 4 && !wide) ? 1 : 0;
    if (c->ph.narrow) c->var.fuse = 0;   // its P3 instance runs unfused
    for (int i = 0; i < 2; ++i)   // captured graphs hold the old arguments
        if (c->gexec[i]) {
            (void)hipGraphExecDestroy(c->gexec[i]);
            c->gexec[i] = nullptr;
        }
    return SHRT_OK;
}

shrt_status shrt_get_variants(shrt_ctx* c, int* p1, int* p2, int* p3, int* fuse) {
    if (!c || !p1 || !p2 || !p3 || !fuse) return fail(SHRT_ERR_INVALID, "null argument");
    *p1 = c->var.p1;
    *p2 = c->var.p2;
    *p3 = c->var.p3;
    *fuse = c->var.fuse;
    return SHRT_OK;
}

shrt_status shrt_set_fusion(shrt_ctx* c, int on) {
    if (!c || on < 0 || on > 1) return fail(SHRT_ERR_INVALID, "bad fusion flag");
    if (on && c->ph.narrow) return fail(SHRT_ERR_INVALID, "the FP32-arithmetic update (shrt_set_update_precision 0) runs unfused");
    c->var.fuse = on;
    for (int i = 0; i < 2; ++i)
        if (c->gexec[i]) {
            (void)hipGraphExecDestroy(c->gexec[i]);
            c->gexec[i] = nullptr;
        }
    return SHRT_OK;
}

shrt_status shrt_kernel_times(shrt_ctx* c, int nsteps, double* ms3) {
    if (!c || !ms3 || nsteps <= 0) return fail(SHRT_ERR_INVALID, "bad arguments");
    std::vector<hipEvent_t> ev((size_t)nsteps * 6);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    hipError_t err = hipSuccess;
    for (int k = 0; k < nsteps && err == hipSuccess; ++k) {
        hipEvent_t* e = &ev[(size_t)k * 6];
        const int par = c->cur;
        auto launch = [&](int which) -> hipError_t {
            const int sp = k == nsteps - 1;
            if (c->prec == 4) {
                const StepArgs<float> a = step_args<float>(c, par, sp);
                return which == 0 ? Launch<float>::p1(a, nullptr, 1, c->var.p1, c->stream)
                     : which == 1 ? Launch<float>::p2(a, nullptr, 1, c->var.p2, c->stream)
                                  : Launch<float>::p3(a, nullptr, 1, c->var.p3, c->stream);
            }
            const StepArgs<double> a = step_args<double>(c, par, sp);
            return which == 0 ? Launch<double>::p1(a, nullptr, 1, c->var.p1, c->stream)
                 : which == 1 ? Launch<double>::p2(a, nullptr, 1, c->var.p2, c->stream)
                              : Launch<double>::p3(a, nullptr, 1, c->var.p3, c->stream);
        };
        // each dispatch records its own start and end (e[2w], e[2w+1])
        for (int w = 0; w < 3 && err == hipSuccess; ++w) {
            g_launch_timing = LaunchTiming{e[2 * w], e[2 * w + 1]};
            err = launch(w);
            g_launch_timing = LaunchTiming{};
        }
        c->cur ^= 1;
    }
    HIP_TRY(err);
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < nsteps; ++k)
        for (int w = 0; w < 3; ++w) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[(size_t)k * 6 + 2 * w], ev[(size_t)k * 6 + 2 * w + 1]));
            acc[w] += ms;
        }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int w = 0; w < 3; ++w) ms3[w] = acc[w] / nsteps;
    return shrt_synchronize(c);
}

shrt_status shrt_fused_kernel_time(shrt_ctx* c, int nsteps, double* ms) {
    if (!c || !ms || nsteps <= 0) return fail(SHRT_ERR_INVALID, "bad arguments");
    if (c->ph.narrow) return fail(SHRT_ERR_INVALID, "the FP32-arithmetic update (shrt_set_update_precision 0) runs unfused");
    // a batch of K = nsteps + 1 steps launched as a fused captured batch would
    // launch it (P1; then P2 and P3+P1 per step; P2 and P3 last), each fused
    // dispatch recording events at its own start and end
    const int K = nsteps + 1;
    std::vector<hipEvent_t> ev((size_t)nsteps * 2);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    hipError_t err = hipSuccess;
    const int par0 = c->cur;
    for (int k = 0; k < K && err == hipSuccess; ++k) {
        const int in_par = par0 ^ (k & 1);
        const bool last = k == K - 1;
        const int sp = k + 1 == K - 1;
        auto go = [&](auto tag) -> hipError_t {
            using T = decltype(tag);
            const StepArgs<T> a = step_args<T>(c, in_par, 0, sp);
            hipError_t e = k == 0 ? Launch<T>::p1(a, nullptr, 1, c->var.p1, c->stream) : hipSuccess;
            if (e == hipSuccess) e = Launch<T>::p2(a, nullptr, 1, c->var.p2, c->stream);
            if (e != hipSuccess) return e;
            if (last) return Launch<T>::p3(a, nullptr, 1, c->var.p3, c->stream);
            g_launch_timing = LaunchTiming{ev[(size_t)k * 2], ev[(size_t)k * 2 + 1]};
            e = Launch<T>::p31(a, nullptr, 1, c->stream);
            g_launch_timing = LaunchTiming{};
            return e;
        };
        err = c->prec == 4 ? go(float{}) : go(double{});
    }
    c->cur = par0 ^ (K & 1);
    HIP_TRY(err);
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc = 0.0;
    for (int k = 0; k < nsteps; ++k) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ev[(size_t)k * 2], ev[(size_t)k * 2 + 1]));
        acc += t;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms = acc / nsteps;
    return shrt_synchronize(c);
}

shrt_status shrt_get_time(shrt_ctx* c, shrt_time* out) {
    if (!c || !out) return fail(SHRT_ERR_INVALID, "null argument");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    out->istep = v.istep;
    out->time = v.time;
    out->delt = v.delt;
    out->remap_count = v.remap_count;
    out->halvings = v.halvings;
    out->last_remap_step = v.last_remap_step;
    return SHRT_OK;
}

shrt_status shrt_set_time(shrt_ctx* c, int64_t istep, double time, double delt) {
    if (!c || istep < 0 || !(delt > 0)) return fail(SHRT_ERR_INVALID, "shrt_set_time: bad arguments");
    if (std::fabs(c->ph.sigma) * delt * 3.0 > 1.0)
        return fail(SHRT_ERR_INVALID, "shrt_set_time: |sigma|*delt must be below 1/3 (remaps would overlap)");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    // Ring slots are addressed by istep mod 3.  The reference's variables keep
    // their buffers when only the counter changes, so when the new count
    // addresses other slots, each logical slot's contents and frame move to
    // the slot the new count gives it (a 3-cycle of copies through one
    // scratch buffer per ring).
    const long long n0 = v.istep;
    if (((n0 - istep) % 3 + 3) % 3 != 0) {
        int po[3], pn[3];
        for (int j = 0; j < 3; ++j) { po[j] = logical_slot(n0, j); pn[j] = logical_slot(istep, j); }
        int src_of[3];                                   // new slot pn[j] takes old slot po[j]
        for (int j = 0; j < 3; ++j) src_of[pn[j]] = po[j];
        const size_t bytes = spec_bytes(c);
        for (int r = 0; r < 4; ++r) {
            void** b = c->ring[r];
            const int d0 = 0, d1 = src_of[d0], d2 = src_of[d1];   // a 3-cycle: src_of[d2] == d0
            HIP_TRY(hipMemcpyAsync(c->tmp_spec, b[d0], bytes, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(b[d0], b[d1], bytes, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(b[d1], b[d2], bytes, hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(b[d2], c->tmp_spec, bytes, hipMemcpyDeviceToDevice, c->stream));
        }
        int ea[3], ed[3];
        for (int k = 0; k < 3; ++k) { ea[k] = v.slot_epoch_a[src_of[k]]; ed[k] = v.slot_epoch_d[src_of[k]]; }
        for (int k = 0; k < 3; ++k) { v.slot_epoch_a[k] = ea[k]; v.slot_epoch_d[k] = ed[k]; }
    }
    v.istep = istep;
    v.time = time;
    v.delt = delt;
    return write_ctl(c, v);
}

shrt_status shrt_get_shift(shrt_ctx* c, double* ky_shift, int* jump) {
    if (!c) return fail(SHRT_ERR_INVALID, "null context");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    std::vector<double> s(c->g.nkx);
    HIP_TRY(hipMemcpyAsync(s.data(), c->s + (size_t)c->cur * c->g.nkx, sizeof(double) * c->g.nkx,
                           hipMemcpyDeviceToHost, c->stream));
    // shearing_ky (shear.cu:280-281) sets jump = round(shift/kx[1]) every step,
    // before a remap zeroes the shift: after a remap step that is the remap's
    // table, otherwise it follows from the current shift; row 0 keeps 0.
    const bool remapped = v.remap_count > 0 && v.last_remap_step == v.istep - 1;
    if (jump && remapped)
        HIP_TRY(hipMemcpyAsync(jump, c->jump + (size_t)(v.epoch & 1) * c->g.nkx, sizeof(int) * c->g.nkx,
                               hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ky_shift) memcpy(ky_shift, s.data(), sizeof(double) * c->g.nkx);
    if (jump && !remapped) {
        jump[0] = 0;
        for (int i = 1; i < c->g.nkx; ++i) jump[i] = v.istep == 0 ? 0 : (int)std::round(s[i] / c->kx_h[1]);
    }
    return SHRT_OK;
}

shrt_status shrt_set_shift(shrt_ctx* c, const double* ky_shift) {
    if (!c || !ky_shift) return fail(SHRT_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpyAsync(c->s + (size_t)c->cur * c->g.nkx, ky_shift, sizeof(double) * c->g.nkx,
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_get_spectral(shrt_ctx* c, int field, void* host) {
    if (!c || !host) return fail(SHRT_ERR_INVALID, "null argument");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const void* src;
    int ep;
    switch (field) {
        case SHRT_OMEGA: src = c->ring[0][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_RHO: src = c->ring[2][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_PHI: src = c->aphi; ep = v.phi_epoch; break;
        default: return fail(SHRT_ERR_INVALID, "unknown spectral field");
    }
    const int* jmp;
    if ((st = view_for(c, v, ep, &jmp)) != SHRT_OK) return st;
    HIP_TRY(view_copy(c, src, c->tmp_spec, jmp));
    HIP_TRY(hipMemcpyAsync(host, c->tmp_spec, spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_get_spectral_modes(shrt_ctx* c, int field, int nx_c, int ny_c, void* host) {
    if (!c || !host) return fail(SHRT_ERR_INVALID, "null argument");
    int sz[8];
    if (shrt_sizes(nx_c, ny_c, sz) != SHRT_OK) return fail(SHRT_ERR_INVALID, "bad coarse grid");
    const int nkx_c = sz[0], nky_c = sz[1], nkxh_c = sz[2];
    if (nkxh_c > c->g.nkxh || nky_c > c->g.nky)
        return fail(SHRT_ERR_INVALID, "shrt_get_spectral_modes: the coarse grid keeps modes this grid does not");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const void* src;
    int ep;
    switch (field) {
        case SHRT_OMEGA: src = c->ring[0][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_RHO: src = c->ring[2][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_PHI: src = c->aphi; ep = v.phi_epoch; break;
        default: return fail(SHRT_ERR_INVALID, "unknown spectral field");
    }
    const int* jmp;
    if ((st = view_for(c, v, ep, &jmp)) != SHRT_OK) return st;
    if (jmp) {
        HIP_TRY(view_copy(c, src, c->tmp_spec, jmp));
        src = c->tmp_spec;
    }
    // coarse row i is fine row i (kx >= 0) or nkx - (nkx_c - i) (kx < 0); its first nky_c modes
    const size_t es = c->esz, pitch_f = (size_t)c->g.nky * es, pitch_c = (size_t)nky_c * es;
    const char* s8 = static_cast<const char*>(src);
    char* h8 = static_cast<char*>(host);
    HIP_TRY(hipMemcpy2DAsync(h8, pitch_c, s8, pitch_f, pitch_c, nkxh_c, hipMemcpyDeviceToHost, c->stream));
    if (nkx_c > nkxh_c)
        HIP_TRY(hipMemcpy2DAsync(h8 + (size_t)nkxh_c * pitch_c, pitch_c,
                                 s8 + (size_t)(c->g.nkx - (nkx_c - nkxh_c)) * pitch_f, pitch_f, pitch_c,
                                 nkx_c - nkxh_c, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_get_spectral_unremapped(shrt_ctx* c, int field, void* host) {
    if (!c || !host) return fail(SHRT_ERR_INVALID, "null argument");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const void* src;
    switch (field) {
        case SHRT_OMEGA: src = c->ring[0][slot]; break;
        case SHRT_RHO: src = c->ring[2][slot]; break;
        case SHRT_PHI: src = c->aphi; break;
        default: return fail(SHRT_ERR_INVALID, "unknown spectral field");
    }
    HIP_TRY(hipMemcpyAsync(host, src, spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_set_spectral(shrt_ctx* c, int field, const void* host) {
    if (!c || !host) return fail(SHRT_ERR_INVALID, "null argument");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    void* dst;
    switch (field) {
        case SHRT_OMEGA: dst = c->ring[0][slot]; break;
        case SHRT_RHO: dst = c->ring[2][slot]; break;
        case SHRT_PHI: dst = c->aphi; break;
        default: return fail(SHRT_ERR_INVALID, "unknown spectral field");
    }
    if (field != SHRT_PHI && (st = materialise(c, v, field == SHRT_OMEGA ? 2 : 0, slot)) != SHRT_OK) return st;
    HIP_TRY(hipMemcpyAsync(dst, host, spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    if (field == SHRT_OMEGA) {
        HIP_TRY(phi_from_omega(c, dst, c->aphi, c->s + (size_t)c->cur * c->g.nkx));
        v.phi_epoch = v.epoch;
    }
    if (field == SHRT_PHI) v.phi_epoch = v.epoch;
    else v.slot_epoch_a[slot] = v.epoch;
    return write_ctl(c, v);
}

shrt_status shrt_get_buffer(shrt_ctx* c, int ring, int j, void* host) {
    if (!c || !host || ring < 0 || ring > 3 || j < 0 || j > 2) return fail(SHRT_ERR_INVALID, "bad arguments");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = logical_slot(v.istep, j);
    const int ep = (ring == 0 || ring == 2) ? v.slot_epoch_a[slot] : v.slot_epoch_d[slot];
    const int* jmp;
    if ((st = view_for(c, v, ep, &jmp)) != SHRT_OK) return st;
    HIP_TRY(view_copy(c, c->ring[ring][slot], c->tmp_spec, jmp));
    HIP_TRY(hipMemcpyAsync(host, c->tmp_spec, spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_set_buffer(shrt_ctx* c, int ring, int j, const void* host) {
    if (!c || !host || ring < 0 || ring > 3 || j < 0 || j > 2) return fail(SHRT_ERR_INVALID, "bad arguments");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = logical_slot(v.istep, j);
    if ((st = materialise(c, v, ring ^ 2, slot)) != SHRT_OK) return st;   // the sibling: a_omg<->a_rho, d_omg<->d_rho
    HIP_TRY(hipMemcpyAsync(c->ring[ring][slot], host, spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    if (ring == 0 || ring == 2) v.slot_epoch_a[slot] = v.epoch;
    else v.slot_epoch_d[slot] = v.epoch;
    return write_ctl(c, v);
}

// The advection term of the current state, vx*d(omega)/dx + vy*d(omega)/dy in
// the sheared frame, as the next step's advect_omg computes it
// (poisson_bracket_shear, shear.cu:395-434): phi = omega/kperp2 with the
// current shift (neg_lapinv_shear, shear.cu:367-374), vx = ktox(i*kys*phi),
// vy = ktox(-i*kx*phi) (get_vector_shear, shear.cu:378-391), and the
// derivatives of omega.  ktox of -i*kx*omega is -d(omega)/dx, hence the negated
// first product.  The two physical factors live in G beyond the nkxh*ny
// elements that ktox_dev uses for its intermediate.
static shrt_status bracket_to_host(shrt_ctx* c, const StepCtl& v, void* host, int lab_frame) {
    if (lab_frame) return fail(SHRT_ERR_INVALID, "the bracket is output in the sheared frame only");
    if (c->ph.linear) return fail(SHRT_ERR_INVALID, "a linear run computes no bracket (time_integral.cu:496-498)");
    const size_t pb = phys_bytes(c), off = (size_t)c->g.I * c->esz;
    if (off + 2 * pb > 6 * (size_t)c->g.I * c->esz) return fail(SHRT_ERR_UNSUPPORTED, "bracket scratch too small for this grid");
    const int slot = slot_a1(v.istep);
    const int* jmp;
    shrt_status st = view_for(c, v, v.slot_epoch_a[slot], &jmp);
    if (st != SHRT_OK) return st;
    const double* s_now = c->s + (size_t)c->cur * c->g.nkx;
    void* f0 = static_cast<char*>(c->G) + off;
    void* f1 = static_cast<char*>(f0) + pb;
    const long long n = (long long)c->g.nx * c->g.ny;
    HIP_TRY(view_copy(c, c->ring[0][slot], c->tmp_spec, jmp));   // omega in the current frame
    HIP_TRY(phi_from_omega(c, c->tmp_spec, c->tmp_spec2, s_now));
    HIP_TRY(ktox_dev(c, c->tmp_spec2, f0, s_now, nullptr, 1));   // vx
    HIP_TRY(ktox_dev(c, c->tmp_spec, f1, s_now, nullptr, 2));    // -d(omega)/dx
    HIP_TRY(mul_acc(c, f0, f1, c->tmp_phys, n, 1, 0));
    HIP_TRY(ktox_dev(c, c->tmp_spec2, f0, s_now, nullptr, 2));   // vy
    HIP_TRY(ktox_dev(c, c->tmp_spec, f1, s_now, nullptr, 1));    // d(omega)/dy
    HIP_TRY(mul_acc(c, f0, f1, c->tmp_phys, n, 0, 1));
    HIP_TRY(hipMemcpyAsync(host, c->tmp_phys, pb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_get_physical(shrt_ctx* c, int pf, int lab_frame, void* host) {
    if (!c || !host) return fail(SHRT_ERR_INVALID, "null argument");
    if (lab_frame && (pf == SHRT_P_VX || pf == SHRT_P_VY))
        return fail(SHRT_ERR_INVALID, "velocities are output in the sheared frame only");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const double* s_now = c->s + (size_t)c->cur * c->g.nkx;
    const double* s_phi = v.istep > 0 ? c->s + (size_t)(c->cur ^ 1) * c->g.nkx : s_now;
    const void* src = nullptr;
    int mode = 0, ep = v.epoch;
    const double* s = s_now;
    switch (pf) {
        case SHRT_P_OMEGA: src = c->ring[0][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_P_RHO: src = c->ring[2][slot]; ep = v.slot_epoch_a[slot]; break;
        case SHRT_P_PHI: src = c->aphi; ep = v.phi_epoch; break;
        case SHRT_P_VX: src = c->aphi; mode = 1; s = s_phi; break;   // raw aphi of the last CFL check
        case SHRT_P_VY: src = c->aphi; mode = 2; s = s_phi; break;
        case SHRT_P_BRACKET: return bracket_to_host(c, v, host, lab_frame);
        default: return fail(SHRT_ERR_INVALID, "unknown physical field");
    }
    const int* jmp = nullptr;
    if (mode == 0 && (st = view_for(c, v, ep, &jmp)) != SHRT_OK) return st;
    // output_fields (file_access.cu:250-259): the lab frame when sigma != 0
    if (lab_frame && c->ph.sigma != 0.0) HIP_TRY(ktox_shear_dev(c, src, c->tmp_phys, s_now, jmp));
    else HIP_TRY(ktox_dev(c, src, c->tmp_phys, s, jmp, mode));
    HIP_TRY(hipMemcpyAsync(host, c->tmp_phys, phys_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_xtok(shrt_ctx* c, const void* phys_in, void* spec_out) {
    if (!c || !phys_in || !spec_out) return fail(SHRT_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpyAsync(c->tmp_phys, phys_in, phys_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(xtok_dev(c, c->tmp_phys, c->tmp_spec));
    HIP_TRY(hipMemcpyAsync(spec_out, c->tmp_spec, spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free xtok on device buffers
// ---------------------------------------------------------------------------
// The x-forward intermediate of shrt_xtok_noalloc (nkxh*ny complex values, every
// element written by the x pass before the y pass reads it) lives in this
// file-scope array instead of the context's H buffer.  Every GPU the library
// runs on holds its own copy until the process ends; calls that overlap on one
// GPU share it and race.  Host code never uses the array's name as an address:
// each device's address comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;   // device -> address of `scratch` on it

// The address of `scratch` on device `dev`, which must be current.
shrt_status scratch_ptr(int dev, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return SHRT_OK;
    }
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        return fail(SHRT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    *out = p;
    return SHRT_OK;
}

template <typename T>
hipError_t xtok_scratch(shrt_ctx* c, const void* phys, void* spec, void* scr) {
    AuxArgs<T> a = aux_args<T>(c, phys, scr, nullptr, nullptr, true, 0);
    hipError_t e = Launch<T>::xfwd(a, c->stream);
    if (e != hipSuccess) return e;
    AuxArgs<T> b = aux_args<T>(c, scr, spec, nullptr, nullptr, false, 0);
    return Launch<T>::yfwd(b, c->stream);
}
}  // namespace

extern "C" {

shrt_status shrt_xtok_noalloc(shrt_ctx* c, const void* phys_dev, void* spec_dev) {
    if (!c || !phys_dev || !spec_dev) return fail(SHRT_ERR_INVALID, "null argument");
    const size_t need = (size_t)c->g.I * c->esz;
    if (need > kNoAllocScratchBytes)
        return fail(SHRT_ERR_INVALID, "shrt_xtok_noalloc: a " + std::to_string(c->g.nx) + "x" +
                                          std::to_string(c->g.ny) + " grid needs " + std::to_string(need) +
                                          " bytes of scratch; the library's holds " +
                                          std::to_string(kNoAllocScratchBytes) +
                                          " (2048x2048 in FP32, 2048x1024 in FP64)");
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    if (prev != c->device) HIP_TRY(hipSetDevice(c->device));
    void* scr = nullptr;
    shrt_status st = scratch_ptr(c->device, &scr);
    hipError_t e = hipSuccess;
    if (st == SHRT_OK)
        e = c->prec == 4 ? xtok_scratch<float>(c, phys_dev, spec_dev, scr)
                         : xtok_scratch<double>(c, phys_dev, spec_dev, scr);
    if (prev != c->device) (void)hipSetDevice(prev);
    if (st != SHRT_OK) return st;
    HIP_TRY(e);
    return SHRT_OK;
}

shrt_status shrt_ktox(shrt_ctx* c, const void* spec_in, void* phys_out) {
    if (!c || !spec_in || !phys_out) return fail(SHRT_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpyAsync(c->tmp_spec, spec_in, spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(ktox_dev(c, c->tmp_spec, c->tmp_phys, c->s + (size_t)c->cur * c->g.nkx, nullptr, 0));
    HIP_TRY(hipMemcpyAsync(phys_out, c->tmp_phys, phys_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_ktox_shear(shrt_ctx* c, const void* spec_in, void* phys_out) {
    if (!c || !spec_in || !phys_out) return fail(SHRT_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpyAsync(c->tmp_spec, spec_in, spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(ktox_shear_dev(c, c->tmp_spec, c->tmp_phys, c->s + (size_t)c->cur * c->g.nkx, nullptr));
    HIP_TRY(hipMemcpyAsync(phys_out, c->tmp_phys, phys_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_shearing_field(shrt_ctx* c, const int* jump, const void* in, void* out) {
    if (!c || !jump || !in || !out) return fail(SHRT_ERR_INVALID, "null argument");
    for (int i = 0; i < c->g.nkx; ++i) {
        const int j = jump[i];
        if (j > c->g.nky + c->g.nkx || j < -(c->g.nky + c->g.nkx))
            return fail(SHRT_ERR_INVALID, "jump out of range");
    }
    HIP_TRY(hipMemcpyAsync(c->jump_tmp, jump, sizeof(int) * c->g.nkx, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tmp_spec, in, spec_bytes(c), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(view_copy(c, c->tmp_spec, c->tmp_spec2, c->jump_tmp));
    HIP_TRY(hipMemcpyAsync(out, c->tmp_spec2, spec_bytes(c), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHRT_OK;
}

shrt_status shrt_diagnostics(shrt_ctx* c, shrt_diag* out) {
    if (!c || !out) return fail(SHRT_ERR_INVALID, "null argument");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const int* jmp;
    if ((st = view_for(c, v, v.slot_epoch_a[slot], &jmp)) != SHRT_OK) return st;
    const double* s = c->s + (size_t)c->cur * c->g.nkx;
    hipError_t e = c->prec == 4
        ? Launch<float>::diag(c->g, c->ring[0][slot], c->ring[2][slot], jmp, c->kx_d, c->ky_d, s, c->diag_partial,
                              kDiagBlocks, c->diag_out, c->stream)
        : Launch<double>::diag(c->g, c->ring[0][slot], c->ring[2][slot], jmp, c->kx_d, c->ky_d, s, c->diag_partial,
                               kDiagBlocks, c->diag_out, c->stream);
    HIP_TRY(e);
    double h[4];
    HIP_TRY(hipMemcpyAsync(h, c->diag_out, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->energy = h[0];
    out->enstrophy = h[1];
    out->rho_var = h[2];
    return SHRT_OK;
}

shrt_status shrt_get_halvings(shrt_ctx* c, shrt_halving* out, int cap, int* count) {
    if (!c || !count || cap < 0) return fail(SHRT_ERR_INVALID, "bad arguments");
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    int n = v.log_count < cap ? v.log_count : cap;
    std::vector<HalvingRec> h(n > 0 ? n : 1);
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(h.data(), c->hlog, sizeof(HalvingRec) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    for (int i = 0; i < n && out; ++i) {
        out[i].istep = h[i].istep;
        out[i].time = h[i].time;
        out[i].value = h[i].value;
        out[i].delt = h[i].delt;
        out[i].which = h[i].which;
        out[i].count = h[i].pad_;
        for (int k = 0; k < 4; ++k) {
            out[i].counts[k] = h[i].counts[k];
            out[i].values[k] = h[i].values[k];
        }
        out[i].delt_before = h[i].delt_before;
    }
    *count = n;
    return SHRT_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: independent instances advanced by one captured graph
// ---------------------------------------------------------------------------
}  // extern "C"

struct shrt_ensemble {
    std::vector<shrt_ctx*> m;
    std::vector<std::vector<int>> groups;                 // members of one precision and shape
    int K = 0;                                            // steps per graph (even)
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    struct Graph {
        hipGraphExec_t exec = nullptr;
        std::vector<void*> bufs;                          // the batches' argument blocks
    };
    std::map<std::vector<int>, Graph> graphs;             // by the members' step parities
};

namespace {

// The argument blocks of one group for every (step parity, aphi flag): member
// i at block [(2*k1 + sp) * n + i] steps from parity par_i ^ k1; sp is
// store_phi, or store_phi_next when the group runs fused steps.
template <typename T>
hipError_t group_args(shrt_ensemble* e, const std::vector<int>& grp, const std::vector<int>& par, void** out) {
    const size_t n = grp.size();
    const bool fuse = e->m[grp[0]]->var.fuse;
    std::vector<StepArgs<T>> h(4 * n);
    for (int k1 = 0; k1 < 2; ++k1)
        for (int sp = 0; sp < 2; ++sp)
            for (size_t i = 0; i < n; ++i)
                h[(2 * k1 + sp) * n + i] = fuse ? step_args<T>(e->m[grp[i]], par[grp[i]] ^ k1, 0, sp)
                                                : step_args<T>(e->m[grp[i]], par[grp[i]] ^ k1, sp, 0);
    hipError_t err = hipMalloc(out, sizeof(StepArgs<T>) * h.size());
    if (err != hipSuccess) return err;
    return hipMemcpy(*out, h.data(), sizeof(StepArgs<T>) * h.size(), hipMemcpyHostToDevice);
}

template <typename T>
hipError_t group_step(shrt_ensemble* e, const std::vector<int>& grp, const void* args, int k, hipStream_t s) {
    const size_t n = grp.size();
    const Variants& v = e->m[grp[0]]->var;   // the group runs its first member's variants
    const bool last = k == e->K - 1;
    const int sp = v.fuse ? (k + 1 == e->K - 1) : last;   // the batch layout of enqueue_batch_step_t
    // a group of one takes its arguments by value
    const StepArgs<T>* d = n == 1 ? nullptr : static_cast<const StepArgs<T>*>(args) + (2 * (k & 1) + sp) * n;
    const StepArgs<T> geom = v.fuse ? step_args<T>(e->m[grp[0]], e->m[grp[0]]->cur ^ (k & 1), 0, sp)
                                    : step_args<T>(e->m[grp[0]], e->m[grp[0]]->cur ^ (k & 1), sp, 0);
    hipError_t err = (!v.fuse || k == 0) ? Launch<T>::p1(geom, d, (int)n, v.p1, s) : hipSuccess;
    if (err == hipSuccess) err = Launch<T>::p2(geom, d, (int)n, v.p2, s);
    if (err == hipSuccess)
        err = (v.fuse && !last) ? Launch<T>::p31(geom, d, (int)n, s) : Launch<T>::p3(geom, d, (int)n, v.p3, s);
    return err;
}

// One graph holding, per group of equal precision and shape, a chain of K
// batched steps (each kernel launched once for all the group's members, the
// member as blockIdx.y); the chains fork from and join into the ensemble's
// stream, so groups run concurrently.
shrt_status ensemble_graph(shrt_ensemble* e, const std::vector<int>& par, hipGraphExec_t* out) {
    auto it = e->graphs.find(par);
    if (it != e->graphs.end()) { *out = it->second.exec; return SHRT_OK; }
    shrt_ensemble::Graph gr;
    for (const auto& grp : e->groups) {
        void* buf = nullptr;
        const hipError_t err = e->m[grp[0]]->prec == 4 ? group_args<float>(e, grp, par, &buf)
                                                       : group_args<double>(e, grp, par, &buf);
        if (buf) gr.bufs.push_back(buf);
        if (err != hipSuccess) {
            for (void* b : gr.bufs) (void)hipFree(b);
            return fail(SHRT_ERR_ALLOC, std::string("ensemble arguments: ") + hipGetErrorString(err));
        }
    }
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipEventRecord(e->fork, e->stream);
    for (size_t gi = 0; gi < e->groups.size() && err == hipSuccess; ++gi) {
        const auto& grp = e->groups[gi];
        hipStream_t s = e->m[grp[0]]->stream;
        err = hipStreamWaitEvent(s, e->fork, 0);
        for (int k = 0; k < e->K && err == hipSuccess; ++k)
            err = e->m[grp[0]]->prec == 4 ? group_step<float>(e, grp, gr.bufs[gi], k, s)
                                          : group_step<double>(e, grp, gr.bufs[gi], k, s);
        if (err == hipSuccess) err = hipEventRecord(e->joins[gi], s);
        if (err == hipSuccess) err = hipStreamWaitEvent(e->stream, e->joins[gi], 0);
    }
    hipError_t e2 = hipStreamEndCapture(e->stream, &graph);
    hipError_t e3 = hipSuccess;
    if (err == hipSuccess && e2 == hipSuccess) e3 = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (err != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        for (void* b : gr.bufs) (void)hipFree(b);
        const hipError_t bad = err != hipSuccess ? err : (e2 != hipSuccess ? e2 : e3);
        return fail(SHRT_ERR_DEVICE, std::string("ensemble graph: ") + hipGetErrorString(bad));
    }
    e->graphs[par] = gr;
    *out = gr.exec;
    return SHRT_OK;
}

}  // namespace

extern "C" {

shrt_status shrt_ensemble_create(shrt_ctx* const* members, int n, int steps_per_graph, shrt_ensemble** out) {
    if (!members || !out || n < 1) return fail(SHRT_ERR_INVALID, "null argument or empty ensemble");
    if (steps_per_graph < 2 || (steps_per_graph & 1)) return fail(SHRT_ERR_INVALID, "steps_per_graph must be even and >= 2");
    for (int i = 0; i < n; ++i) {
        if (!members[i]) return fail(SHRT_ERR_INVALID, "null member");
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) return fail(SHRT_ERR_INVALID, "a context appears twice");
    }
    auto* e = new shrt_ensemble;
    e->m.assign(members, members + n);
    e->K = steps_per_graph;
    for (int i = 0; i < n; ++i) {   // group by precision and shape
        const shrt_ctx* c = members[i];
        size_t gi = 0;
        for (; gi < e->groups.size(); ++gi) {
            const shrt_ctx* r = members[e->groups[gi][0]];
            if (r->prec == c->prec && r->g.nx == c->g.nx && r->g.ny == c->g.ny && r->ph.linear == c->ph.linear &&
                r->ph.narrow == c->ph.narrow)
                break;
        }
        if (gi == e->groups.size()) e->groups.emplace_back();
        e->groups[gi].push_back(i);
    }
    e->joins.assign(e->groups.size(), nullptr);
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (size_t i = 0; ok && i < e->joins.size(); ++i)
        ok = hipEventCreateWithFlags(&e->joins[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        shrt_ensemble_destroy(e);
        return fail(SHRT_ERR_DEVICE, "ensemble stream or event creation failed");
    }
    *out = e;
    return SHRT_OK;
}

shrt_status shrt_ensemble_destroy(shrt_ensemble* e) {
    if (!e) return SHRT_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto& kv : e->graphs) {
        (void)hipGraphExecDestroy(kv.second.exec);
        for (void* b : kv.second.bufs) (void)hipFree(b);
    }
    for (hipEvent_t ev : e->joins)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return SHRT_OK;
}

shrt_status shrt_ensemble_step(shrt_ensemble* e, int nsteps) {
    if (!e || nsteps < 0) return fail(SHRT_ERR_INVALID, "shrt_ensemble_step: bad arguments");
    const size_t n = e->m.size();
    int full = nsteps / e->K;
    if (full > 0) {
        std::vector<int> par(n);
        for (size_t i = 0; i < n; ++i) par[i] = e->m[i]->cur;
        hipGraphExec_t ex = nullptr;
        shrt_status st = ensemble_graph(e, par, &ex);
        if (st != SHRT_OK) return st;
        // the graph runs after each member's earlier work, and before its later work
        for (size_t i = 0; i < n; ++i) {
            HIP_TRY(hipEventRecord(e->fork, e->m[i]->stream));
            HIP_TRY(hipStreamWaitEvent(e->stream, e->fork, 0));
        }
        for (int b = 0; b < full; ++b) HIP_TRY(hipGraphLaunch(ex, e->stream));   // K even: parities unchanged
        HIP_TRY(hipEventRecord(e->fork, e->stream));
        for (size_t i = 0; i < n; ++i) HIP_TRY(hipStreamWaitEvent(e->m[i]->stream, e->fork, 0));
    }
    const int rest = nsteps - full * e->K;
    for (size_t i = 0; i < n && rest > 0; ++i) {
        shrt_status st = shrt_step_async(e->m[i], rest);
        if (st != SHRT_OK) return st;
    }
    return SHRT_OK;
}

shrt_status shrt_ensemble_synchronize(shrt_ensemble* e) {
    if (!e) return fail(SHRT_ERR_INVALID, "null ensemble");
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (shrt_ctx* c : e->m) {
        shrt_status st = shrt_synchronize(c);
        if (st != SHRT_OK) return st;
    }
    return SHRT_OK;
}

shrt_status shrt_ensemble_time_steps(shrt_ensemble* e, int nsteps, double* ms_total) {
    if (!e || !ms_total || nsteps < e->K || nsteps % e->K) return fail(SHRT_ERR_INVALID, "bad arguments");
    std::vector<int> par(e->m.size());
    for (size_t i = 0; i < e->m.size(); ++i) par[i] = e->m[i]->cur;
    hipGraphExec_t ex = nullptr;
    shrt_status st = ensemble_graph(e, par, &ex);   // built outside the timed region
    if (st != SHRT_OK) return st;
    st = shrt_ensemble_synchronize(e);
    if (st != SHRT_OK) return st;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, e->stream));
    st = shrt_ensemble_step(e, nsteps);
    HIP_TRY(hipEventRecord(e1, e->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != SHRT_OK) return st;
    *ms_total = ms;
    return shrt_ensemble_synchronize(e);
}

}  // extern "C"

namespace {

template <typename T>
shrt_status closure_correct_t(shrt_closure* net, shrt_ctx* c, const double* scale_in, double sigma,
                              const double* scale_out, double dt_tot, double guard, int kcut, int* applied) {
    StepCtl v;
    shrt_status st = read_ctl(c, &v);
    if (st != SHRT_OK) return st;
    const int slot = slot_a1(v.istep);
    const double* s_now = c->s + (size_t)c->cur * c->g.nkx;
    const long long NP = (long long)c->g.nx * c->g.ny, M = (long long)c->g.nkx * c->g.nky;
    float* x = closure_input(net);
    const float* out = closure_output(net);
    // inputs: omega, rho, phi in physical space (phi as stored, through its view) / scale, then sigma
    const void* src[3] = {c->ring[0][slot], c->ring[2][slot], c->aphi};
    const int ep[3] = {v.slot_epoch_a[slot], v.slot_epoch_a[slot], v.phi_epoch};
    for (int i = 0; i < 3; ++i) {
        const int* jmp;
        if ((st = view_for(c, v, ep[i], &jmp)) != SHRT_OK) return st;
        HIP_TRY(ktox_dev(c, src[i], c->tmp_phys, s_now, jmp, 0));
        HIP_TRY(Launch<T>::net_in(c->tmp_phys, x + i * NP, NP, scale_in[i], 0.0, c->stream));
    }
    HIP_TRY(Launch<T>::net_in(c->tmp_phys, x + 3 * NP, NP, 0.0, sigma, c->stream));
    HIP_TRY(closure_forward_on(net, c->stream));
    // corrections: xtok of out * scale_out, kept in G beyond ktox_dev's intermediate
    void* corr[2] = {static_cast<char*>(c->G) + (size_t)c->g.I * c->esz,
                     static_cast<char*>(c->G) + ((size_t)c->g.I + (size_t)M) * c->esz};
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(Launch<T>::net_out(out + i * NP, c->tmp_phys, NP, scale_out[i], c->stream));
        HIP_TRY(xtok_dev(c, c->tmp_phys, corr[i]));
        if (kcut > 0) HIP_TRY(Launch<T>::lowpass(c->g, corr[i], kcut, c->stream));
    }
    // the state in the current frame, norms and the finiteness of the output
    void* cur[2] = {c->tmp_spec, c->tmp_spec2};
    for (int i = 0; i < 2; ++i) {
        const int* jmp;
        if ((st = view_for(c, v, v.slot_epoch_a[slot], &jmp)) != SHRT_OK) return st;
        HIP_TRY(view_copy(c, c->ring[2 * i][slot], cur[i], jmp));
    }
    double* red = c->diag_partial;   // the diagnostics' partial sums: free outside shrt_diagnostics
    HIP_TRY(Launch<T>::sumsq(corr[0], nullptr, M, red + 0, c->stream));
    HIP_TRY(Launch<T>::sumsq(corr[1], nullptr, M, red + 1, c->stream));
    HIP_TRY(Launch<T>::sumsq(cur[0], nullptr, M, red + 2, c->stream));
    HIP_TRY(Launch<T>::sumsq(cur[1], nullptr, M, red + 3, c->stream));
    HIP_TRY(Launch<T>::sumsq(nullptr, out, 2 * NP, red + 4, c->stream));
    double r[5];
    HIP_TRY(hipMemcpyAsync(r, red, sizeof r, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the fallback: a non-finite output or a correction larger than guard x the field
    const bool ok = r[4] == 0.0 && std::isfinite(r[0]) && std::isfinite(r[1]) &&
                    !(std::sqrt(r[0]) * dt_tot > guard * std::sqrt(r[2])) &&
                    !(std::sqrt(r[1]) * dt_tot > guard * std::sqrt(r[3]));
    *applied = ok ? 1 : 0;
    if (!ok) return SHRT_OK;
    for (int i = 0; i < 2; ++i)
        HIP_TRY(Launch<T>::axpy_round(cur[i], corr[i], dt_tot, c->ring[2 * i][slot], M, c->stream));
    // as shrt_set_spectral does for omega: phi of the new omega, both slots in the current frame
    HIP_TRY(phi_from_omega(c, c->ring[0][slot], c->aphi, s_now));
    v.phi_epoch = v.epoch;
    v.slot_epoch_a[slot] = v.epoch;
    return write_ctl(c, v);
}

}  // namespace

extern "C" {

shrt_status shrt_closure_correct(shrt_closure* net, shrt_ctx* c, const double* scale_in, double sigma,
                                 const double* scale_out, double dt_tot, double guard, int kcut, int* applied) {
    if (!net || !c || !scale_in || !scale_out || !applied) return fail(SHRT_ERR_INVALID, "null argument");
    *applied = 0;
    if (closure_grid(net) != c->g.nx || c->g.nx != c->g.ny)
        return fail(SHRT_ERR_INVALID, "the network's grid differs from the context's");
    for (int i = 0; i < 3; ++i)
        if (!(scale_in[i] > 0.0)) return fail(SHRT_ERR_INVALID, "input scales must be positive");
    if (!std::isfinite(dt_tot) || !std::isfinite(guard) || !std::isfinite(sigma))
        return fail(SHRT_ERR_INVALID, "non-finite dt, guard or sigma");
    if (2 * (size_t)c->g.nkx * c->g.nky > 5 * (size_t)c->g.I)
        return fail(SHRT_ERR_UNSUPPORTED, "correction scratch too small for this grid");
    return c->prec == 4 ? closure_correct_t<float>(net, c, scale_in, sigma, scale_out, dt_tot, guard, kcut, applied)
                        : closure_correct_t<double>(net, c, scale_in, sigma, scale_out, dt_tot, guard, kcut, applied);
}

}  // extern "C"

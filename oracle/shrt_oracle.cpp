// shrt_oracle.cpp -- FP64 CPU restatement of shearing-rt's pseudospectral step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP path
// and the timed CPU baseline; nothing in the product links it.  It follows the
// reference's fields/fourier/shear/time_integral path function by function (the
// reference file:line is given above each function) in double precision, with
// its own FFT (the reference calls cuFFT, which is absent here), built with
// -ffp-contract=off so every multiply and add rounds as written.
//
// Layouts are the reference's own:
//   spectral  f[ikx*nky + iky]   (kx order 0..nkxh-1, then -(nkxh-1)..-1)
//   physical  f[ix*ny + iy]      (y contiguous)
//
// Deliberate, documented departures from the reference (the reference code is
// racy or undefined there):
//   * shearing_ky: the reset test is taken once after the update of every row
//     (the reference's unsynchronised test can miss rows when nkx-1 > 1024).
//   * the remap fold reads index jump-iky only when it is < nky (the reference
//     would read out of bounds, which only happens for nx > ny).
//   * pl_max_search: a plain max over all nx*ny values (the reference reads past
//     the end when nx*ny is not a multiple of 2*nthread).
// Parity status of every function is listed in the oracle section of the
// project design notes; all are pinned by tests/test_oracle_*.py.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct C { double x, y; };

// cmplx.h:34-139 -- the operators the hot path uses, in the same rounding order.
inline C cadd(C a, C b) { return {a.x + b.x, a.y + b.y}; }
inline C csub(C a, C b) { return {a.x - b.x, a.y - b.y}; }
inline C cmul(C a, C b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }  // cuCmul
inline C rmul(double a, C b) { return {a * b.x, a * b.y}; }
inline C cmulr(C a, double b) { return {a.x * b, a.y * b}; }
inline C cdivr(C a, double b) { return {a.x / b, a.y / b}; }
inline C cneg(C a) { return {-a.x, -a.y}; }
const C CIMAG = {0.0, 1.0};
const C CZERO = {0.0, 0.0};

// ---------------------------------------------------------------------------
// FFT (stands in for cuFFT, fft.cu:111-119).  Unnormalised, forward = e^{-i}.
// Power-of-two lengths use an iterative radix-2 transform; other lengths use a
// direct O(n^2) DFT.  Twiddles come from long-double cos/sin, rounded once.
// ---------------------------------------------------------------------------
struct FFTPlan {
    int n = 0;
    bool pow2 = false;
    std::vector<C> w;        // w[k] = exp(-2 pi i k / n)
    std::vector<int> rev;
};

FFTPlan make_plan(int n) {
    FFTPlan p;
    p.n = n;
    p.pow2 = n > 0 && (n & (n - 1)) == 0;
    p.w.resize(n);
    for (int k = 0; k < n; ++k) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        p.w[k] = {(double)cosl(ang), (double)sinl(ang)};
    }
    if (p.pow2) {
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        p.rev.resize(n);
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            p.rev[i] = r;
        }
    }
    return p;
}

// dir = -1 forward, +1 inverse.  `a` has stride 1, length p.n.  `tmp` >= n.
void fft_exec(const FFTPlan& p, C* a, int dir, C* tmp) {
    const int n = p.n;
    if (n <= 1) return;
    if (!p.pow2) {
        for (int k = 0; k < n; ++k) {
            C acc = CZERO;
            for (int j = 0; j < n; ++j) {
                int idx = (int)(((long long)k * j) % n);
                C w = p.w[idx];
                if (dir > 0) w.y = -w.y;
                acc = cadd(acc, cmul(a[j], w));
            }
            tmp[k] = acc;
        }
        memcpy(a, tmp, sizeof(C) * n);
        return;
    }
    for (int i = 0; i < n; ++i)
        if (i < p.rev[i]) { C t = a[i]; a[i] = a[p.rev[i]]; a[p.rev[i]] = t; }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; ++j) {
                C w = p.w[j * step];
                if (dir > 0) w.y = -w.y;
                C u = a[i + j];
                C v = cmul(a[i + j + half], w);
                a[i + j] = cadd(u, v);
                a[i + j + half] = csub(u, v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The oracle's state: the globals of fields.cu:13-25, fft.cu:7-8,
// fourier.cu:14-17, shear.cu:21-34, time_integral.cu:25-36.
// ---------------------------------------------------------------------------
struct Orc {
    // fft.cu:7
    int nx, ny, nkx, nky, nkxh, nkxh2, nkxpad, ncy;
    // fields.cu:13-16, time_integral.cu:25-31
    double delt, cfl_num, nu, kappa, sigma, g, rho0, rho0_prime, rho_eps1, rho_eps2;
    double Lx, Ly, dx, dy, eps;
    bool noise_flag, linear_flag;
    // fourier.cu:14
    std::vector<double> kx, ky, kperp2;
    // shear.cu:21-23
    std::vector<int> ikx_indexed, jump;
    std::vector<double> ky_shift, kperp2_shear;
    int jump_flag;
    // fields.cu:18-25 (spectral state; ring slots rotate by swapping vectors)
    std::vector<C> aomg[3], domg[3], arho[3], drho[3], aphi;
    std::vector<C> kdata_bef[3];   // omega, phi, rho as update_shear hands them to k_data_bef at the last remap
    std::vector<double> xx, yy, vx, vy, rtmp, ddx_r, ddy_r;
    // bookkeeping added by the restatement
    int istep;
    double time;
    int remap_count;
    int last_remap_step;
    int halvings;
    FFTPlan px, py;
    int nthreads;
};

int orc_threads(const Orc* o) { return o->nthreads > 0 ? o->nthreads : 1; }

// ---------------------------------------------------------------------------
// fft.cu:136-166  xtok: R2C(nx,ny), truncate to nkx x nky, divide by nx*ny.
// ---------------------------------------------------------------------------
void xtok(Orc* o, const double* in, C* out) {
    const int nx = o->nx, ny = o->ny, nky = o->nky, nkx = o->nkx;
    std::vector<C> work((size_t)nx * nky);
    // y transform of every row (R2C keeps k < ncy; only k < nky survive the truncation)
    #pragma omp parallel num_threads(orc_threads(o))
    {
        std::vector<C> row(ny), tmp(ny);
        #pragma omp for schedule(static)
        for (int ix = 0; ix < nx; ++ix) {
            for (int iy = 0; iy < ny; ++iy) row[iy] = {in[(size_t)ix * ny + iy], 0.0};
            fft_exec(o->py, row.data(), -1, tmp.data());
            for (int k = 0; k < nky; ++k) work[(size_t)ix * nky + k] = row[k];
        }
    }
    // x transform of every retained ky column, then scale_dealias (fft.cu:155-166)
    const double scale = (double)(nx * ny);
    #pragma omp parallel num_threads(orc_threads(o))
    {
        std::vector<C> col(nx), tmp(nx);
        #pragma omp for schedule(static)
        for (int k = 0; k < nky; ++k) {
            for (int ix = 0; ix < nx; ++ix) col[ix] = work[(size_t)ix * nky + k];
            fft_exec(o->px, col.data(), -1, tmp.data());
            for (int xid = 0; xid < nkx; ++xid) {
                int src = xid < o->nkxh ? xid : xid + o->nkxpad;
                out[(size_t)xid * nky + k] = cdivr(col[src], scale);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// fft.cu:170-204  ktox: pad2d into nx x ncy, then C2R (unnormalised).  The 2-D
// C2R is taken x first (C2C along x for every ky column), then C2R along y
// using the real part of the ky=0 (and Nyquist) terms -- numpy's irfft2 order.
// ---------------------------------------------------------------------------
void ktox(Orc* o, const C* in, double* out) {
    const int nx = o->nx, ny = o->ny, nky = o->nky, ncy = o->ncy;
    std::vector<C> work((size_t)nx * nky);
    #pragma omp parallel num_threads(orc_threads(o))
    {
        std::vector<C> col(nx), tmp(nx);
        #pragma omp for schedule(static)
        for (int k = 0; k < nky; ++k) {
            // pad2d (fft.cu:189-204)
            for (int xid = 0; xid < nx; ++xid) {
                if (xid < o->nkxh) col[xid] = in[(size_t)xid * nky + k];
                else if (xid > o->nkxh2) col[xid] = in[(size_t)(xid - o->nkxpad) * nky + k];
                else col[xid] = CZERO;
            }
            fft_exec(o->px, col.data(), +1, tmp.data());
            for (int ix = 0; ix < nx; ++ix) work[(size_t)ix * nky + k] = col[ix];
        }
    }
    #pragma omp parallel num_threads(orc_threads(o))
    {
        std::vector<C> row(ny), tmp(ny);
        #pragma omp for schedule(static)
        for (int ix = 0; ix < nx; ++ix) {
            for (int iy = 0; iy < ny; ++iy) row[iy] = CZERO;
            for (int k = 0; k < nky && k < ncy; ++k) {
                C v = work[(size_t)ix * nky + k];
                bool selfconj = (k == 0) || (2 * k == ny);
                if (selfconj) { row[k] = {v.x, 0.0}; continue; }
                row[k] = v;
                row[ny - k] = {v.x, -v.y};
            }
            fft_exec(o->py, row.data(), +1, tmp.data());
            for (int iy = 0; iy < ny; ++iy) out[(size_t)ix * ny + iy] = row[iy].x;
        }
    }
}

// ---------------------------------------------------------------------------
// shear.cu:467-527 ktox_shear, the lab-frame output transform:
//   idft_shear_y (:481-513): for every stored kx row a direct one-sided sum over
//     ky >= 0 with weight 1 for ky = 0 and 2 otherwise, at the sheared
//     wavenumber ky + ky_shift[kx], in the reference's operation order;
//   pad2d (:515-527): kx rows into nx rows, zero between nkxh and nkxh2;
//   ktox_1d (fft.cu:208-250): C2C inverse along x per y, keep the real part.
// ---------------------------------------------------------------------------
void ktox_shear(Orc* o, const C* in, double* out) {
    const int nx = o->nx, ny = o->ny, nkx = o->nkx, nky = o->nky;
    std::vector<C> ctmp2((size_t)nkx * ny);
    #pragma omp parallel for num_threads(orc_threads(o)) schedule(static)
    for (int xid = 0; xid < nkx; ++xid) {
        const double ky_shift = o->ky_shift[xid];
        for (int yid = 0; yid < ny; ++yid) {
            const double y_data = o->yy[yid];
            const C* f = in + (size_t)xid * nky;
            double ky = o->ky[0] + ky_shift;
            C out_data;
            out_data.x = f[0].x * cos(ky * y_data) - f[0].y * sin(ky * y_data);
            out_data.y = f[0].y * cos(ky * y_data) + f[0].x * sin(ky * y_data);
            for (int iky = 1; iky < nky; ++iky) {
                ky = o->ky[iky] + ky_shift;
                out_data.x += 2 * f[iky].x * cos(ky * y_data) - 2 * f[iky].y * sin(ky * y_data);
                out_data.y += 2 * f[iky].y * cos(ky * y_data) + 2 * f[iky].x * sin(ky * y_data);
            }
            ctmp2[(size_t)xid * ny + yid] = out_data;
        }
    }
    #pragma omp parallel num_threads(orc_threads(o))
    {
        std::vector<C> col(nx), tmp(nx);
        #pragma omp for schedule(static)
        for (int yid = 0; yid < ny; ++yid) {
            for (int xid = 0; xid < nx; ++xid) {   // pad2d
                if (xid < o->nkxh) col[xid] = ctmp2[(size_t)xid * ny + yid];
                else if (xid > o->nkxh2) col[xid] = ctmp2[(size_t)(xid - o->nkxpad) * ny + yid];
                else col[xid] = CZERO;
            }
            fft_exec(o->px, col.data(), +1, tmp.data());
            for (int xid = 0; xid < nx; ++xid) out[(size_t)xid * ny + yid] = col[xid].x;
        }
    }
}

// ---------------------------------------------------------------------------
// fourier.cu:170-240 and shear.cu:326-374 -- pointwise spectral operators.
// ---------------------------------------------------------------------------
void negative(Orc* o, double* f) {                      // fourier.cu:170-176
    const size_t n = (size_t)o->nx * o->ny;
    for (size_t i = 0; i < n; ++i) f[i] = -f[i];
}
void ddx(Orc* o, const C* in, C* out) {                 // fourier.cu:180-188
    for (int t = 0; t < o->nkx * o->nky; ++t)
        out[t] = cmul(cmulr(CIMAG, o->kx[t / o->nky]), in[t]);
}
void neg_ddx(Orc* o, const C* in, C* out) {             // fourier.cu:190-198
    for (int t = 0; t < o->nkx * o->nky; ++t)
        out[t] = cmul(cmulr(cneg(CIMAG), o->kx[t / o->nky]), in[t]);
}
void ddy(Orc* o, const C* in, C* out) {                 // fourier.cu:202-210
    for (int t = 0; t < o->nkx * o->nky; ++t)
        out[t] = cmul(cmulr(CIMAG, o->ky[t % o->nky]), in[t]);
}
void neg_lapinv(Orc* o, const C* in, C* out) {          // fourier.cu:233-240
    for (int t = 1; t <= o->nkx * o->nky - 1; ++t) out[t] = cdivr(in[t], o->kperp2[t]);
}
void get_kperp2_shear(Orc* o) {                         // shear.cu:326-338
    for (int t = 0; t < o->nkx * o->nky; ++t) {
        int xid = t / o->nky, yid = t % o->nky;
        double kxs = o->kx[xid];
        double kys = o->ky[yid] + o->ky_shift[xid];
        o->kperp2_shear[t] = kxs * kxs + kys * kys;
    }
}
void ddy_shear(Orc* o, const C* in, C* out) {           // shear.cu:342-354
    for (int t = 0; t < o->nkx * o->nky; ++t) {
        int xid = t / o->nky, yid = t % o->nky;
        double kys = o->ky[yid] + o->ky_shift[xid];
        out[t] = cmul(cmulr(CIMAG, kys), in[t]);
    }
}
void neg_lapinv_shear(Orc* o, const C* in, C* out) {    // shear.cu:367-374
    for (int t = 1; t <= o->nkx * o->nky - 1; ++t) out[t] = cdivr(in[t], o->kperp2_shear[t]);
}

void get_vector(Orc* o, const C* aphi, double* vx, double* vy) {   // fourier.cu:244-257
    std::vector<C> tmp((size_t)o->nkx * o->nky);
    ddy(o, aphi, tmp.data());
    ktox(o, tmp.data(), vx);
    neg_ddx(o, aphi, tmp.data());
    ktox(o, tmp.data(), vy);
}
void get_vector_shear(Orc* o, const C* aphi, double* vx, double* vy) {   // shear.cu:378-391
    std::vector<C> tmp((size_t)o->nkx * o->nky);
    ddy_shear(o, aphi, tmp.data());
    ktox(o, tmp.data(), vx);
    neg_ddx(o, aphi, tmp.data());
    ktox(o, tmp.data(), vy);
}

// shear.cu:395-434  poisson_bracket_shear + calc_poisson_bracket.  The reference's
// product kernel reads the global velocity arrays, which are the ones passed in.
void poisson_bracket_shear(Orc* o, const double* vx, const double* vy, const C* in, double* out) {
    std::vector<C> tmp((size_t)o->nkx * o->nky);
    ddx(o, in, tmp.data());
    ktox(o, tmp.data(), o->ddx_r.data());
    ddy_shear(o, in, tmp.data());
    ktox(o, tmp.data(), o->ddy_r.data());
    const size_t n = (size_t)o->nx * o->ny;
    for (size_t t = 0; t < n; ++t) out[t] = vx[t] * o->ddx_r[t] + vy[t] * o->ddy_r[t];
}

// time_integral.cu:386-409  add_v0 + maxvalue_search (max over |f| of all points)
double maxabs(Orc* o, const double* f) {
    double m = 0;
    const size_t n = (size_t)o->nx * o->ny;
    for (size_t i = 0; i < n; ++i) { double v = fabs(f[i]); if (m < v) m = v; }
    return m;
}

// time_integral.cu:345-384  check_cfl
void check_cfl(Orc* o) {
    neg_lapinv_shear(o, o->aomg[0].data(), o->aphi.data());
    get_vector_shear(o, o->aphi.data(), o->vx.data(), o->vy.data());
    const int nx = o->nx, ny = o->ny;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            o->rtmp[(size_t)ix * ny + iy] = o->vx[(size_t)ix * ny + iy] + o->sigma * o->yy[iy];
    double cfl_vx = maxabs(o, o->rtmp.data());
    while ((cfl_vx * o->delt / o->dx) > o->cfl_num) { o->delt /= 2.0; o->halvings++; }
    double cfl_vy = maxabs(o, o->vy.data());
    while ((cfl_vy * o->delt / o->dy) > o->cfl_num) { o->delt /= 2.0; o->halvings++; }
    double cfl_grho0 = o->g / o->rho0;
    while ((cfl_grho0 * o->delt / o->dx) > o->cfl_num) { o->delt /= 2.0; o->halvings++; }
    double cfl_rho0prime = o->rho0_prime;
    while ((cfl_rho0prime * o->delt / o->dx) > o->cfl_num) { o->delt /= 2.0; o->halvings++; }
}

// time_integral.cu:474-484  renew_fields: (a0,a1,a2) <- (a2,a0,a1); aphi is not rotated.
void renew_fields(Orc* o) {
    std::vector<C>* rings[4] = {o->aomg, o->domg, o->arho, o->drho};
    for (auto r : rings) {
        std::swap(r[2], r[1]);   // r2 <- old r1
        std::swap(r[1], r[0]);   // r1 <- old r0, r0 <- old r2
    }
}

// time_integral.cu:621-695  explicit kernels
void eulerf(Orc* o, C* a0, const C* a1, const C* d0) {
    for (int t = 0; t < o->nkx * o->nky; ++t) a0[t] = cadd(a1[t], rmul(o->delt, d0[t]));
}
void ab2_plain(Orc* o, C* a0, const C* a1, const C* d0, const C* d1) {
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cadd(a1[t], rmul(o->delt, csub(rmul(1.5, d0[t]), rmul(0.5, d1[t]))));
}
void ab2_bdf(Orc* o, C* a0, const C* a1, const C* a2, const C* d0, const C* d1) {
    const double c = (2.0 / 3.0) * o->delt;
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cadd(cdivr(csub(rmul(4.0, a1[t]), a2[t]), 3.0),
                     rmul(c, csub(rmul(2.0, d0[t]), d1[t])));
}
void ab3_plain(Orc* o, C* a0, const C* a1, const C* d0, const C* d1, const C* d2) {
    const double c = o->delt / 12.0;
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cadd(a1[t], rmul(c, cadd(csub(rmul(23.0, d0[t]), rmul(16.0, d1[t])), rmul(5.0, d2[t]))));
}
void ab3_bdf(Orc* o, C* a0, const C* a1, const C* a2, const C* d0, const C* d1, const C* d2) {
    const double c = o->delt / 9.0;
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cadd(cdivr(csub(rmul(4.0, a1[t]), a2[t]), 3.0),
                     rmul(c, cadd(csub(rmul(16.0, d0[t]), rmul(14.0, d1[t])), rmul(4.0, d2[t]))));
}
// time_integral.cu:747-767  implicit diffusion with the current kperp2_shear
void eulerb(Orc* o, C* a0, double diss) {
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cdivr(a0[t], 1.0 + o->delt * diss * o->kperp2_shear[t]);
}
void bdf2(Orc* o, C* a0, double diss) {
    for (int t = 0; t < o->nkx * o->nky; ++t)
        a0[t] = cdivr(a0[t], 1.0 + (2.0 / 3.0) * o->delt * diss * o->kperp2_shear[t]);
}

// time_integral.cu:545-554, 608-617  buoyancy and background-gradient terms
void add_ddxrho(Orc* o, C* d0, const C* arho1) {
    const double c = o->g / o->rho0;
    for (int t = 0; t < o->nkx * o->nky; ++t)
        d0[t] = csub(d0[t], cmul(cmulr(rmul(c, CIMAG), o->kx[t / o->nky]), arho1[t]));
}
void add_ddxphi(Orc* o, C* d0, const C* aphi) {
    for (int t = 0; t < o->nkx * o->nky; ++t)
        d0[t] = cadd(d0[t], cmul(cmulr(rmul(o->rho0_prime, CIMAG), o->kx[t / o->nky]), aphi[t]));
}

void explicit_update(Orc* o, std::vector<C>* a, std::vector<C>* d, double diff) {
    // time_integral.cu:512-541 / 575-605
    switch (o->istep) {
        case 0: eulerf(o, a[0].data(), a[1].data(), d[0].data()); break;
        case 1:
            if (diff < o->eps) ab2_plain(o, a[0].data(), a[1].data(), d[0].data(), d[1].data());
            else ab2_bdf(o, a[0].data(), a[1].data(), a[2].data(), d[0].data(), d[1].data());
            break;
        default:
            if (diff < o->eps) ab3_plain(o, a[0].data(), a[1].data(), d[0].data(), d[1].data(), d[2].data());
            else ab3_bdf(o, a[0].data(), a[1].data(), a[2].data(), d[0].data(), d[1].data(), d[2].data());
            break;
    }
}
void dissipate(Orc* o, std::vector<C>& a0, double diff) {
    // time_integral.cu:699-743
    if (diff < o->eps) return;
    if (o->istep == 0) eulerb(o, a0.data(), diff);
    else bdf2(o, a0.data(), diff);
}

// time_integral.cu:488-543
void advect_omg(Orc* o) {
    const int M = o->nkx * o->nky;
    if (o->linear_flag) {
        for (int t = 0; t < M; ++t) o->domg[0][t] = CZERO;
    } else {
        poisson_bracket_shear(o, o->vx.data(), o->vy.data(), o->aomg[1].data(), o->rtmp.data());
        negative(o, o->rtmp.data());
        xtok(o, o->rtmp.data(), o->domg[0].data());
    }
    add_ddxrho(o, o->domg[0].data(), o->arho[1].data());
    explicit_update(o, o->aomg, o->domg, o->nu);
}
// time_integral.cu:558-606
void advect_rho(Orc* o) {
    const int M = o->nkx * o->nky;
    if (o->linear_flag) {
        for (int t = 0; t < M; ++t) o->drho[0][t] = CZERO;
    } else {
        poisson_bracket_shear(o, o->vx.data(), o->vy.data(), o->arho[1].data(), o->rtmp.data());
        negative(o, o->rtmp.data());
        xtok(o, o->rtmp.data(), o->drho[0].data());
    }
    add_ddxphi(o, o->drho[0].data(), o->aphi.data());
    explicit_update(o, o->arho, o->drho, o->kappa);
}

// shear.cu:275-294  shearing_ky, with the reset test taken once after every row
// has been advanced (the intended, race-free reading of the reference kernel).
// Returns 1 when the remap fires.
int shearing_ky(Orc* o) {
    const int nkx = o->nkx;
    for (int t = 1; t <= nkx - 1; ++t) {
        o->ky_shift[t] = o->ky_shift[t] - o->sigma * o->kx[t] * o->delt;
        o->jump[t] = (int)round(o->ky_shift[t] / o->kx[1]);
    }
    if (nkx > 1 && fabs(o->ky_shift[1] - (-o->kx[1])) < 1e-10) {
        for (int t = 1; t <= nkx - 1; ++t) o->ky_shift[t] = 0;
        return 1;
    }
    return 0;
}

// shear.cu:296-322  shearing_field: out[x,y] = in[x, y-jump[x]] (or 0 past nky) for
// jump<=0; for jump>0 the folded part is conj(in[ikx_indexed[x], jump-y]).
void shearing_field(const Orc* o, const int* jmp, const C* in, C* out) {
    const int nkx = o->nkx, nky = o->nky;
    for (int t = 0; t < nkx * nky; ++t) {
        int xid = t / nky, yid = t % nky;
        int j = jmp[xid];
        if (j <= 0) {
            out[t] = (yid - j < nky) ? in[xid * nky + (yid - j)] : CZERO;
        } else {
            if (yid - j > 0) {
                out[t] = in[xid * nky + (yid - j)];
            } else {
                int src = j - yid;
                if (src < nky) {
                    C v = in[o->ikx_indexed[xid] * nky + src];
                    out[t] = {v.x, -v.y};
                } else {
                    out[t] = CZERO;
                }
            }
        }
    }
}

// shear.cu:211-273  update_shear
void update_shear(Orc* o) {
    o->jump_flag = shearing_ky(o);
    if (o->jump_flag) {
        // k_data_bef(time+delt, istep+1) (shear.cu:236): the fields before the remap
        o->kdata_bef[0] = o->aomg[0];
        o->kdata_bef[1] = o->aphi;
        o->kdata_bef[2] = o->arho[0];
        std::vector<C> tmp((size_t)o->nkx * o->nky);
        std::vector<C>* rings[4] = {o->aomg, o->domg, o->arho, o->drho};
        for (auto r : rings)
            for (int s = 0; s < 3; ++s) {
                tmp = r[s];
                shearing_field(o, o->jump.data(), tmp.data(), r[s].data());
            }
        tmp = o->aphi;
        shearing_field(o, o->jump.data(), tmp.data(), o->aphi.data());
        o->remap_count++;
        o->last_remap_step = o->istep;
    }
    get_kperp2_shear(o);
}

// time_integral.cu:324-341  time_advance
void time_advance(Orc* o) {
    check_cfl(o);
    renew_fields(o);
    advect_omg(o);
    dissipate(o, o->aomg[0], o->nu);
    advect_rho(o);
    dissipate(o, o->arho[0], o->kappa);
    update_shear(o);
    o->istep++;
    o->time += o->delt;
}

// time_integral.cu:229-259, 261-320  initialize + init_dis + disturb (glibc rand()).
void initialize_reference(Orc* o) {
    const int nx = o->nx, ny = o->ny;
    std::vector<double> omgz((size_t)nx * ny), rho((size_t)nx * ny);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            omgz[(size_t)ix * ny + iy] = 0;
            rho[(size_t)ix * ny + iy] = o->rho_eps1 * cos(2 * o->xx[ix] + 10 * o->yy[iy]);
        }
    xtok(o, omgz.data(), o->aomg[0].data());
    xtok(o, rho.data(), o->arho[0].data());
    if (o->noise_flag) {
        const int nkx = o->nkx, nky = o->nky;
        srand(2000);                                   // time_integral.cu:18, 274
        for (int ikx = 0; ikx < nkx; ++ikx)
            for (int iky = 0; iky < nky; ++iky) {
                double f1 = 0, f2 = 0;
                if (fabs(o->kx[ikx]) <= 2 && fabs(o->ky[iky]) != 0 && fabs(o->ky[iky]) <= 10) {
                    f1 = o->rho_eps2 * (2 * ((double)rand() / (1.0 + RAND_MAX)) - 1.0);
                    f2 = o->rho_eps2 * (2 * ((double)rand() / (1.0 + RAND_MAX)) - 1.0);
                }
                C& r = o->arho[0][(size_t)ikx * nky + iky];
                r.x = r.x + f1;
                r.y = r.y + f2;
            }
    }
    neg_lapinv(o, o->aomg[0].data(), o->aphi.data());
    get_vector(o, o->aphi.data(), o->vx.data(), o->vy.data());
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

// ===========================================================================
// C ABI used by tests/ and bench.py (ctypes).
// ===========================================================================
extern "C" {

struct orc_params {
    int nx, ny;
    double Lx, Ly, delt, cfl_num;
    double nu, kappa, sigma, g, rho0, rho0_prime, rho_eps1;
    int noise, linear;
};

void orc_sizes(int nx, int ny, int* out8) {
    // fft.cu:91-96
    int nkx = (nx - 1) / 3 * 2 + 1;
    int nky = (ny - 1) / 3 + 1;
    int nkxh = (nkx + 1) / 2;
    out8[0] = nkx; out8[1] = nky; out8[2] = nkxh; out8[3] = nx - nkxh;
    out8[4] = nx - nkxh * 2 + 1; out8[5] = ny / 2 + 1; out8[6] = nkx * nky; out8[7] = nx * nky;
}

void orc_fft(int n, int dir, double* data /* 2n doubles, interleaved */) {
    FFTPlan p = make_plan(n);
    std::vector<C> tmp(n);
    fft_exec(p, reinterpret_cast<C*>(data), dir, tmp.data());
}

void* orc_create(const orc_params* prm) {
    Orc* o = new Orc();
    o->nx = prm->nx; o->ny = prm->ny;
    int s[8]; orc_sizes(o->nx, o->ny, s);
    o->nkx = s[0]; o->nky = s[1]; o->nkxh = s[2]; o->nkxh2 = s[3]; o->nkxpad = s[4]; o->ncy = s[5];
    o->Lx = prm->Lx; o->Ly = prm->Ly; o->delt = prm->delt; o->cfl_num = prm->cfl_num;
    o->nu = prm->nu; o->kappa = prm->kappa; o->sigma = prm->sigma; o->g = prm->g;
    o->rho0 = prm->rho0; o->rho0_prime = prm->rho0_prime; o->rho_eps1 = prm->rho_eps1;
    o->rho_eps2 = 0.1 * prm->rho_eps1;                 // file_access.cu:164
    o->noise_flag = prm->noise != 0; o->linear_flag = prm->linear != 0;
    o->eps = 1e-10;                                    // time_integral.cu:208
    o->dx = o->Lx / o->nx; o->dy = o->Ly / o->ny;      // main.cu:24
    o->px = make_plan(o->nx); o->py = make_plan(o->ny);
    o->nthreads = 1;
    const int nkx = o->nkx, nky = o->nky, M = nkx * nky;
    // fourier.cu:123-134
    o->kx.assign(nkx, 0.0);
    for (int ikx = 0; ikx < o->nkxh; ikx++) o->kx[ikx] = ikx * 2 * M_PI / o->Lx;
    for (int ikx = 0; ikx < o->nkxh - 1; ikx++) o->kx[nkx - ikx - 1] = -o->kx[ikx + 1];
    o->ky.assign(nky, 0.0);
    for (int iky = 0; iky < nky; iky++) o->ky[iky] = iky * 2 * M_PI / o->Ly;
    o->kperp2.assign(M, 0.0);
    for (int ikx = 0; ikx < nkx; ikx++)
        for (int iky = 0; iky < nky; iky++)
            o->kperp2[ikx * nky + iky] = o->kx[ikx] * o->kx[ikx] + o->ky[iky] * o->ky[iky];
    // shear.cu:149-185
    o->kperp2_shear = o->kperp2;
    o->ikx_indexed.assign(nkx, 0);
    for (int ikx = 1; ikx < nkx; ikx++) o->ikx_indexed[ikx] = nkx - ikx;
    o->jump.assign(nkx, 0);
    o->ky_shift.assign(nkx, 0.0);
    o->jump_flag = 0;
    // fields.cu:120-166 (all spectral buffers zeroed)
    for (int r = 0; r < 3; ++r) {
        o->aomg[r].assign(M, CZERO); o->domg[r].assign(M, CZERO);
        o->arho[r].assign(M, CZERO); o->drho[r].assign(M, CZERO);
    }
    o->aphi.assign(M, CZERO);
    const size_t P = (size_t)o->nx * o->ny;
    o->vx.assign(P, 0.0); o->vy.assign(P, 0.0); o->rtmp.assign(P, 0.0);
    o->ddx_r.assign(P, 0.0); o->ddy_r.assign(P, 0.0);
    // time_integral.cu:236-237
    o->xx.resize(o->nx + 1); o->yy.resize(o->ny + 1);
    for (int ix = 0; ix <= o->nx; ix++) o->xx[ix] = ix * o->dx;
    for (int iy = 0; iy <= o->ny; iy++) o->yy[iy] = iy * o->dy;
    o->istep = 0; o->time = 0.0; o->remap_count = 0; o->last_remap_step = -1; o->halvings = 0;
    return o;
}

void orc_destroy(void* h) { delete static_cast<Orc*>(h); }
void orc_set_threads(void* h, int n) { static_cast<Orc*>(h)->nthreads = n; }

void orc_initialize_reference(void* h) { initialize_reference(static_cast<Orc*>(h)); }

// Seeded initial condition (new; no reference counterpart): omega = 0 and, on every
// retained mode with kmin <= |k| <= kmax, rho_hat = amp*|k|^expo*exp(i*theta) with
// theta = 2*pi*u, u = splitmix64(seed ^ (mode+1)*golden) / 2^64 (53 bits).  The ky=0
// column is made Hermitian (rho_hat(-kx,0) = conj(rho_hat(kx,0))); mode 0 is zero.
void orc_seeded_spectrum(int nx, int ny, double Lx, double Ly, uint64_t seed, double kmin,
                         double kmax, double amp, double expo, double* out /* 2*M */) {
    int s[8]; orc_sizes(nx, ny, s);
    const int nkx = s[0], nky = s[1], nkxh = s[2];
    C* o = reinterpret_cast<C*>(out);
    for (int t = 0; t < nkx * nky; ++t) o[t] = CZERO;
    for (int ikx = 0; ikx < nkx; ++ikx) {
        double kxv = (ikx < nkxh ? ikx : ikx - nkx) * 2 * M_PI / Lx;
        for (int iky = 0; iky < nky; ++iky) {
            if (iky == 0 && ikx >= nkxh) continue;   // filled from the +kx partner below
            double kyv = iky * 2 * M_PI / Ly;
            double km = sqrt(kxv * kxv + kyv * kyv);
            if (km < kmin || km > kmax || km == 0.0) continue;
            uint64_t z = splitmix64(seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(ikx * nky + iky + 1)));
            double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            double th = 2 * M_PI * u;
            double a = amp * pow(km, expo);
            C v = {a * cos(th), a * sin(th)};
            o[ikx * nky + iky] = v;
            if (iky == 0 && ikx > 0) o[(nkx - ikx) * nky] = {v.x, -v.y};
        }
    }
}

// Set a spectral field in the reference layout.  field: 0=omega, 1=rho (slot a0),
// 2=phi.  Setting omega or rho also recomputes aphi = omega/kperp2_shear as
// initialize() does (time_integral.cu:257).
void orc_set_spectral(void* h, int field, const double* in) {
    Orc* o = static_cast<Orc*>(h);
    const C* c = reinterpret_cast<const C*>(in);
    const int M = o->nkx * o->nky;
    std::vector<C>* dst = field == 0 ? &o->aomg[0] : field == 1 ? &o->arho[0] : &o->aphi;
    for (int t = 0; t < M; ++t) (*dst)[t] = c[t];
    if (field == 0) neg_lapinv_shear(o, o->aomg[0].data(), o->aphi.data());
}

// Read a buffer: field 0..3 = aomg0..2 / domg0..2 / arho0..2 / drho0..2 selected by
// (ring, slot); ring 4 = aphi.
void orc_get_buffer(void* h, int ring, int slot, double* out) {
    Orc* o = static_cast<Orc*>(h);
    const std::vector<C>* src =
        ring == 0 ? &o->aomg[slot] : ring == 1 ? &o->domg[slot] : ring == 2 ? &o->arho[slot]
        : ring == 3 ? &o->drho[slot] : &o->aphi;
    memcpy(out, src->data(), sizeof(C) * src->size());
}
void orc_set_buffer(void* h, int ring, int slot, const double* in) {
    Orc* o = static_cast<Orc*>(h);
    std::vector<C>* dst =
        ring == 0 ? &o->aomg[slot] : ring == 1 ? &o->domg[slot] : ring == 2 ? &o->arho[slot]
        : ring == 3 ? &o->drho[slot] : &o->aphi;
    memcpy(dst->data(), in, sizeof(C) * dst->size());
}

// The fields k_data_bef wrote at the last remap (0 omega, 1 phi, 2 rho); zeros
// before the first remap.
void orc_get_kdata_bef(void* h, int field, double* out) {
    Orc* o = static_cast<Orc*>(h);
    const std::vector<C>& v = o->kdata_bef[field];
    if (v.empty()) memset(out, 0, sizeof(C) * (size_t)o->nkx * o->nky);
    else memcpy(out, v.data(), sizeof(C) * v.size());
}

void orc_get_physical(void* h, int which, double* out) {
    Orc* o = static_cast<Orc*>(h);
    const std::vector<double>& v = which == 0 ? o->vx : which == 1 ? o->vy : o->rtmp;
    memcpy(out, v.data(), sizeof(double) * v.size());
}

void orc_step(void* h, int nsteps) {
    Orc* o = static_cast<Orc*>(h);
    for (int i = 0; i < nsteps; ++i) time_advance(o);
}

void orc_set_time(void* h, int istep, double time, double delt) {
    Orc* o = static_cast<Orc*>(h);
    o->istep = istep; o->time = time; o->delt = delt;
}

// state: istep, time, delt, remap_count, last_remap_step, halvings
void orc_get_state(void* h, double* out6) {
    Orc* o = static_cast<Orc*>(h);
    out6[0] = o->istep; out6[1] = o->time; out6[2] = o->delt;
    out6[3] = o->remap_count; out6[4] = o->last_remap_step; out6[5] = o->halvings;
}
void orc_get_shift(void* h, double* ky_shift, int* jump) {
    Orc* o = static_cast<Orc*>(h);
    memcpy(ky_shift, o->ky_shift.data(), sizeof(double) * o->nkx);
    memcpy(jump, o->jump.data(), sizeof(int) * o->nkx);
}
void orc_set_shift(void* h, const double* ky_shift) {
    Orc* o = static_cast<Orc*>(h);
    memcpy(o->ky_shift.data(), ky_shift, sizeof(double) * o->nkx);
    get_kperp2_shear(o);
}

void orc_xtok(void* h, const double* in, double* out) {
    xtok(static_cast<Orc*>(h), in, reinterpret_cast<C*>(out));
}
void orc_ktox(void* h, const double* in, double* out) {
    ktox(static_cast<Orc*>(h), reinterpret_cast<const C*>(in), out);
}
// ktox_shear (shear.cu:467-478) with the oracle's current ky_shift
void orc_ktox_shear(void* h, const double* in, double* out) {
    ktox_shear(static_cast<Orc*>(h), reinterpret_cast<const C*>(in), out);
}
void orc_shearing_field(void* h, const int* jmp, const double* in, double* out) {
    shearing_field(static_cast<Orc*>(h), jmp, reinterpret_cast<const C*>(in), reinterpret_cast<C*>(out));
}
// poisson_bracket_shear of an arbitrary field with the oracle's current vx, vy
void orc_poisson_bracket(void* h, const double* in, double* out) {
    Orc* o = static_cast<Orc*>(h);
    poisson_bracket_shear(o, o->vx.data(), o->vy.data(), reinterpret_cast<const C*>(in), out);
}
// check_cfl alone: recomputes aphi, vx, vy from aomg0 and halves delt as needed.
void orc_check_cfl(void* h) { check_cfl(static_cast<Orc*>(h)); }

// Energy/enstrophy diagnostics in FP64 (new; no reference counterpart):
// Hermitian weights w = 1 (iky = 0), 2 (iky > 0) over the stored half plane.
//   out[0] = E = 1/2 sum w kperp2_shear |phi_hat|^2  with phi_hat = omega/kperp2_shear
//   out[1] = Z = 1/2 sum w |omega_hat|^2
//   out[2] = P = 1/2 sum w |rho_hat|^2
void orc_diagnostics(void* h, double* out3) {
    Orc* o = static_cast<Orc*>(h);
    const int nkx = o->nkx, nky = o->nky;
    double E = 0, Z = 0, P = 0;
    for (int ikx = 0; ikx < nkx; ++ikx)
        for (int iky = 0; iky < nky; ++iky) {
            int t = ikx * nky + iky;
            double w = iky == 0 ? 1.0 : 2.0;
            C om = o->aomg[0][t], rh = o->arho[0][t];
            double k2 = o->kperp2_shear[t];
            if (t > 0) {
                C ph = cdivr(om, k2);
                E += w * k2 * (ph.x * ph.x + ph.y * ph.y);
            }
            Z += w * (om.x * om.x + om.y * om.y);
            P += w * (rh.x * rh.x + rh.y * rh.y);
        }
    out3[0] = 0.5 * E; out3[1] = 0.5 * Z; out3[2] = 0.5 * P;
}

}  // extern "C"

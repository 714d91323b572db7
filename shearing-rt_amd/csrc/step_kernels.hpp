// step_kernels.hpp -- the three fused kernels of one pseudospectral step.
//
// The reference's time_advance (time_integral.cu:324-341) runs ~50 launches,
// 8 cuFFT 2-D transforms and 3 blocking device->host reads per step.  Here the
// step is three kernels and never talks to the host:
//
//   P1  per kx>=0 row pair (kx, -kx), y-inverse transforms.  Reads omega_hat^n,
//       rho_hat^n (through the remap view when a remap is pending), applies the
//       Poisson solve phi = omega/kperp2_shear (shear.cu:367-374) and the ik
//       multipliers (fourier.cu:180-198, shear.cu:342-354), and writes six
//       y-transformed fields G: X(i kys phi), X(-i kx phi), X(i kx w),
//       X(i kys w), X(i kx r), X(i kys r).
//   P2  per y column, x-inverse transforms (two real fields per complex FFT),
//       the brackets vx*fx + vy*fy of omega and rho (shear.cu:415-434), the CFL
//       maxima max|vx + sigma*y| and max|vy| (time_integral.cu:359-371), and
//       the x-forward transform of (-B_omega, -B_rho) (fourier.cu:170-176).
//   P3  per kx>=0 row pair, y-forward transform, truncation and 1/(nx*ny)
//       scaling (fft.cu:155-166), buoyancy (time_integral.cu:545-554, 608-617),
//       Euler / AB2 / AB3 in BDF2 form (:621-695) and implicit diffusion
//       (:747-767); workgroup 0 also does the step's bookkeeping: CFL halving,
//       FP64 shift advance, remap trigger and jump table (shear.cu:275-294).
#pragma once

#include "step_common.hpp"

namespace shrt {

// The argument block of a step kernel: the kernel argument itself, or, for a
// batch of instances of one shape, entry blockIdx.y of a device array.
template <bool BATCH, typename T>
__device__ __forceinline__ const StepArgs<T>& step_args_of(const StepArgs<T>& a1, const StepArgs<T>* args) {
    if constexpr (BATCH) return args[blockIdx.y];
    else return a1;
}

// ---------------------------------------------------------------------------
// P1
// ---------------------------------------------------------------------------
// Sheared wavenumbers of one mode in the working precision.  FP64 keeps the
// reference's operation order (shear.cu:335-336); FP32 rounds ky + s once and
// forms kperp2 in FP32.
template <typename T>
struct ModeK {
    T kx, kys, k2;
};

template <typename T>
__device__ __forceinline__ ModeK<T> mode_k(const double* __restrict__ ky, double s_row, double kxv, int iky) {
    const double kysd = add_rn(ky[iky], s_row);
    ModeK<T> m;
    m.kx = (T)kxv;
    m.kys = (T)kysd;
    if constexpr (sizeof(T) == 8) m.k2 = (T)k2_of(kxv, kysd);
    else m.k2 = m.kx * m.kx + m.kys * m.kys;
    return m;
}

// Multiplier of output O of group GRP: group 0 takes phi = omega/kperp2
// (neg_lapinv_shear, shear.cu:367-374; phi = 0 at the origin) to vx = i kys phi
// (ddy_shear, shear.cu:342-354) and vy = -i kx phi (neg_ddx, fourier.cu:190-198);
// groups 1 and 2 take omega or rho to i kx f (ddx, fourier.cu:180-188) and i kys f.
template <int GRP, int O, typename T>
__device__ __forceinline__ cplx<T> p1_mult(const ModeK<T>& k, bool origin, cplx<T> F) {
    if constexpr (GRP == 0) {
        const cplx<T> phi = origin ? czero<T>() : cdivr(F, k.k2);
        if constexpr (O == 0) return {-k.kys * phi.y, k.kys * phi.x};
        else return {k.kx * phi.y, -k.kx * phi.x};
    } else {
        if constexpr (O == 0) return {-k.kx * F.y, k.kx * F.x};
        else return {-k.kys * F.y, k.kys * F.x};
    }
}

// G stores the P1 rows by position: position p holds row p + 1 for p < nkxh - 1
// and row 0 (kx = 0) at p = nkxh - 1.  Full tiles of kTile positions then hold
// rows 1..nkxh-1 except for a remainder that shares a few workgroups with row
// 0, whose kx-multiplied outputs vanish: at 2048^2 the grid is exactly 1024
// workgroups instead of 1026 (four times the 256 CUs).
__host__ __device__ inline int p1_row_of(int p, int nkxh) { return p == nkxh - 1 ? 0 : p + 1; }
__host__ __device__ inline int g_pos_of(int kx, int nkxh) { return kx == 0 ? nkxh - 1 : kx - 1; }

struct P1Grid {
    int nout;      // outputs per row: 6, or 2 in the linear mode
    int nfull;     // full tiles (rows 1 .. 4*nfull)
    int rem;       // rows left over, row 0 included (1..kTile)
    int nspec;     // transforms of the leftover rows (row 0 needs only vx, d/dy omega, d/dy rho)
    int nspec_wg;  // workgroups that run them, one transform per team
    int grid;
    __host__ __device__ P1Grid(int nkxh, int linear) {
        nout = linear ? 2 : 6;
        nfull = (nkxh - 1) / kTile;
        rem = nkxh - nfull * kTile;
        nspec = nout * (rem - 1) + (linear ? 1 : 3);
        nspec_wg = (nspec + kTile - 1) / kTile;
        grid = nfull * nout + nspec_wg;
    }
};

// The outputs of row 0 that are transformed (the others are zero).
__host__ __device__ inline int p1_row0_output(int j) { return j == 0 ? 0 : (j == 1 ? 3 : 5); }

// y-spectrum of output OI for the row pair (row, -row) in natural order,
// branch-free: element e takes row kx at iky = e (e < nky) or the conjugate of
// row -kx at iky = NY - e (e > NY - nky); padding lanes load an in-bounds
// element and are zeroed.  Group GRP = OI/2 (0: phi -> vx, vy; 1: omega;
// 2: rho), output O = OI%2.
template <int OI, typename T, int NY>
__device__ __forceinline__ void p1_spectrum(const StepArgs<T>& a, int row, bool act, int tid, cplx<T>* v) {
    constexpr int GRP = OI / 2, O = OI % 2;
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT;
    const Geo& g = a.g;
    const int rowc = act ? row : 0;
    const int rn = rowc == 0 ? 0 : g.nkx - rowc;
    const StepCtl* c = a.ctl_in;
    const int s1 = slot_a1(c->istep);
    const cplx<T>* src = GRP == 2 ? a.a_rho[s1] : a.a_omg[s1];
    const int* jmp = (c->slot_epoch_a[s1] != c->epoch) ? a.jump + (c->epoch & 1) * g.nkx : nullptr;
    const double kxp = a.kx[rowc], kxn = a.kx[rn];
    const double sp = a.s_in[rowc], sn = a.s_in[rn];
    constexpr int NKY = NY / 3 + 1;   // == nky for the supported (power-of-two) lengths
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        const bool pos = e < NKY;
        const bool keep = act && (pos || e > NY - NKY);
        const int iky = keep ? (pos ? e : NY - e) : 0;
        const int r = pos ? rowc : rn;
        const cplx<T> F = view_load(src, r, iky, g, jmp);
        const ModeK<T> k = mode_k<T>(a.ky, pos ? sp : sn, pos ? kxp : kxn, iky);
        const cplx<T> z = p1_mult<GRP, O, T>(k, r == 0 && iky == 0, F);
        v[m] = keep ? (pos ? z : cplx<T>{z.x, -z.y}) : czero<T>();
    }
    if (act && tid == 0) {
        // Hermitian part of the ky=0 column, as the x-first 2-D C2R uses it.
        const ModeK<T> k = mode_k<T>(a.ky, sn, kxn, 0);
        const cplx<T> vn = p1_mult<GRP, O, T>(k, rn == 0, view_load(src, rn, 0, g, jmp));
        v[0] = {(v[0].x + vn.x) * T(0.5), (v[0].y - vn.y) * T(0.5)};
    }
}

// aphi = omega/kperp2_shear (neg_lapinv_shear, shear.cu:367-374) for the row
// pair (row, -row), by one team.
template <typename T, int NY>
__device__ __forceinline__ void p1_store_phi(const StepArgs<T>& a, int row, int tid) {
    constexpr int TT = TeamShape<T, NY>::TT;
    const Geo& g = a.g;
    const StepCtl* c = a.ctl_in;
    const int s1 = slot_a1(c->istep);
    const cplx<T>* src = a.a_omg[s1];
    const int* jmp = (c->slot_epoch_a[s1] != c->epoch) ? a.jump + (c->epoch & 1) * g.nkx : nullptr;
    const int rn = row == 0 ? 0 : g.nkx - row;
    for (int iky = tid; iky < g.nky; iky += TT) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (side == 1 && row == 0) break;
            const int r = side ? rn : row;
            const ModeK<T> k = mode_k<T>(a.ky, a.s_in[r], a.kx[r], iky);
            const cplx<T> F = view_load(src, r, iky, g, jmp);
            a.aphi[(long long)r * g.nky + iky] = (r == 0 && iky == 0) ? czero<T>() : cdivr(F, k.k2);
        }
    }
}

// Workgroups of P1.  Items q < nfull*nout: full tile q / nout (positions
// 4t..4t+3, rows 4t+1..4t+4), output q % nout, one row per team, written to G
// through [team][y] staging in 4x4 blocks [t][position][c] (y = 4t + c).
// Items beyond: the leftover rows, one (row, output) transform per team,
// written by the team directly; the first of them also writes the vanishing
// outputs of row 0.
template <typename T, int NY, bool BATCH>
__global__ void __launch_bounds__((kTile * TeamShape<T, NY>::TT), (waves_per_eu<T, NY>()))
k_p1(StepArgs<T> a1, const StepArgs<T>* __restrict__ args) {
    const StepArgs<T>& a = step_args_of<BATCH>(a1, args);
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    constexpr int LP = stage_stride<NY>();
    __shared__ cplx<T> sm[kTile * L];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const P1Grid pg(g.nkxh, a.ph.linear);
    // the outputs of one tile read the same rows: keep them on one XCD
    const int q = xcd_item(blockIdx.x, gridDim.x);
    cplx<T>* lds = sm + team * L;
    const TeamSync<S::TT> sync{};
    cplx<T> v[P];

    if (q < pg.nfull * pg.nout) {
        const int tile = q / pg.nout;
        const int oi = q - tile * pg.nout;
        const int row = tile * kTile + team + 1;
        static_for<0, 6>([&](auto oc) {
            constexpr int OI = decltype(oc)::value;
            if (OI == oi) p1_spectrum<OI, T, NY>(a, row, true, tid, v);
        });
        TeamFFT<T, NY, P, S::RMAX, +1>::run(v, lds, a.twy, tid, sync);
        teams_join<TT>();
#pragma unroll
        for (int m = 0; m < P; ++m) sm[team * LP + tid + m * TT] = v[m];
        __syncthreads();
        cplx<T>* Gf = a.G + (long long)oi * g.I;
#pragma unroll
        for (int m = 0; m < P; ++m) {   // NY*4 staged elements over kTile*TT threads
            const int idx = threadIdx.x + m * kTile * TT;
            const int t = idx >> 4, w = idx & 15;
            Gf[((long long)t * g.nkxh + tile * kTile) * 4 + w] = sm[(w >> 2) * LP + 4 * t + (w & 3)];
        }
        if (oi == 0 && a.store_phi) p1_store_phi<T, NY>(a, row, tid);
        return;
    }

    // leftover rows: item i of this team
    const int j = q - pg.nfull * pg.nout;
    const int i = j * kTile + team;
    const bool act = i < pg.nspec;
    const int nreg = pg.nout * (pg.rem - 1);
    const int pos = i < nreg ? pg.nfull * kTile + i / pg.nout : g.nkxh - 1;
    const int oi = i < nreg ? i % pg.nout : p1_row0_output(i - nreg);
    const int row = p1_row_of(pos, g.nkxh);
    static_for<0, 6>([&](auto oc) {
        constexpr int OI = decltype(oc)::value;
        if (OI == oi) p1_spectrum<OI, T, NY>(a, row, act, tid, v);
    });
    TeamFFT<T, NY, P, S::RMAX, +1>::run(v, lds, a.twy, tid, sync);
    if (act) {
        cplx<T>* Gf = a.G + (long long)oi * g.I;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int y = tid + m * TT;
            Gf[((long long)(y >> 2) * g.nkxh + pos) * 4 + (y & 3)] = v[m];
        }
        if (oi == 0 && a.store_phi) p1_store_phi<T, NY>(a, row, tid);
    }
    if (j == 0) {
        // outputs of row 0 that carry a factor kx = 0
        const int p0 = g.nkxh - 1;
        const int nz = a.ph.linear ? 1 : 3;
        for (int idx = threadIdx.x; idx < nz * NY; idx += blockDim.x) {
            const int z = idx / NY, y = idx - z * NY;
            const int o = z == 0 ? 1 : (z == 1 ? 2 : 4);
            a.G[(long long)o * g.I + ((long long)(y >> 2) * g.nkxh + p0) * 4 + (y & 3)] = czero<T>();
        }
    }
}

// P1 with one (row, output) transform per workgroup of one team: eight
// workgroups per CU that start and finish independently.  Item q < (nkxh-1)*nout
// is row q/nout + 1 (G position q/nout), output q % nout; the last items are
// row 0's transformed outputs, the first of which also writes its vanishing
// ones.  Each team writes its G position directly, in runs of 4 (32 bytes) per
// t; the outputs of neighbouring rows are dispatched together to one XCD, whose
// L2 merges the runs into whole lines.
template <typename T, int NY, bool BATCH>
__global__ void __launch_bounds__((TeamShape<T, NY>::TT), (waves_per_eu<T, NY, TeamShape<T, NY>::TT, 1>()))
k_p1_row(StepArgs<T> a1, const StepArgs<T>* __restrict__ args) {
    const StepArgs<T>& a = step_args_of<BATCH>(a1, args);
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    __shared__ cplx<T> sm[L];
    const Geo& g = a.g;
    const int tid = threadIdx.x;
    const int nout = a.ph.linear ? 2 : 6;
    const int nreg = (g.nkxh - 1) * nout;
    const int q = xcd_item(blockIdx.x, gridDim.x);
    const int pos = q < nreg ? q / nout : g.nkxh - 1;
    const int oi = q < nreg ? q - pos * nout : p1_row0_output(q - nreg);
    const int row = p1_row_of(pos, g.nkxh);
    const TeamSync<S::TT> sync{};
    cplx<T> v[P];
    static_for<0, 6>([&](auto oc) {
        constexpr int OI = decltype(oc)::value;
        if (OI == oi) p1_spectrum<OI, T, NY>(a, row, true, tid, v);
    });
    TeamFFT<T, NY, P, S::RMAX, +1>::run(v, sm, a.twy, tid, sync);
    cplx<T>* Gf = a.G + (long long)oi * g.I;
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int y = tid + m * TT;
        Gf[((long long)(y >> 2) * g.nkxh + pos) * 4 + (y & 3)] = v[m];
    }
    if (oi == 0 && a.store_phi) p1_store_phi<T, NY>(a, row, tid);
    if (q == nreg) {
        // outputs of row 0 that carry a factor kx = 0
        const int nz = a.ph.linear ? 1 : 3;
        for (int idx = tid; idx < nz * NY; idx += TT) {
            const int z = idx / NY, y = idx - z * NY;
            const int o = z == 0 ? 1 : (z == 1 ? 2 : 4);
            a.G[(long long)o * g.I + ((long long)(y >> 2) * g.nkxh + pos) * 4 + (y & 3)] = czero<T>();
        }
    }
}

// The y-inverse transforms of the outputs in the bit set omask of the row pair
// (row, -row) from its spectrum before multipliers: base[m] for element
// e = tid + m*TT (row kx at iky = e, or row -kx at iky = NY - e, not yet
// conjugated; zero in the padding), base_n0 = the mirror row's ky = 0 value,
// with the shifts sn[] and wavenumbers kxr[] of the two rows; written to G at
// the row's position, and aphi = omega/kperp2_shear (shear.cu:367-374) when
// store_phi.  One transform per output in a rolled loop whose lane index is
// opaque, so the per-mode wavenumbers are recomputed per output rather than
// hoisted and kept live (48 more registers) across the transforms.  Output oi
// multiplies X = phi (oi < 2) or the field by i*c: c = kys (0, 3, 5), -kx (1),
// kx (2, 4) -- p1_mult in selected form.
template <typename T, int NY>
__device__ __forceinline__ void p1_outputs(const StepArgs<T>& a, int row, const cplx<T>* base, cplx<T> base_n0,
                                           const double* sn, const double* kxr, unsigned omask,
                                           int store_phi, cplx<T>* lds, int tid) {
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT;
    constexpr int NKY = NY / 3 + 1;
    const Geo& g = a.g;
    const int rn = row == 0 ? 0 : g.nkx - row;
    const int gp = g_pos_of(row, g.nkxh);
    const TeamSync<S::TT> sync{};
#pragma nounroll
    for (unsigned mk = omask; mk != 0; mk &= mk - 1) {
        const int oi = __builtin_ctz(mk);
        int tl = tid;
        asm volatile("" : "+v"(tl));
        tl &= TT - 1;
        const bool use_phi = oi < 2, use_kys = oi == 0 || oi == 3 || oi == 5, neg = oi == 1;
        auto mult = [&](const ModeK<T>& k, bool origin, cplx<T> F) -> cplx<T> {
            if (use_phi) F = origin ? czero<T>() : cdivr(F, k.k2);
            const T cf = use_kys ? k.kys : (neg ? -k.kx : k.kx);
            return {-cf * F.y, cf * F.x};
        };
        cplx<T> v[P];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int e = tl + m * TT;
            const bool pos = e < NKY;
            const bool keep = pos || e > NY - NKY;
            const int iky = keep ? (pos ? e : NY - e) : 0;
            const int side = pos ? 0 : 1;
            const int r = side ? rn : row;
            const ModeK<T> k = mode_k<T>(a.ky, sn[side], kxr[side], iky);
            const cplx<T> z = mult(k, r == 0 && iky == 0, base[m]);
            v[m] = keep ? (pos ? z : cplx<T>{z.x, -z.y}) : czero<T>();
        }
        if (tl == 0) {   // Hermitian part of the ky = 0 column
            const ModeK<T> k = mode_k<T>(a.ky, sn[1], kxr[1], 0);
            const cplx<T> vn = mult(k, rn == 0, base_n0);
            v[0] = {(v[0].x + vn.x) * T(0.5), (v[0].y - vn.y) * T(0.5)};
        }
        if (oi == 0 && store_phi) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                const int e = tl + m * TT;
                const bool pos = e < NKY;
                if (!(pos || e > NY - NKY) || (!pos && row == 0)) continue;
                const int iky = pos ? e : NY - e;
                const int side = pos ? 0 : 1;
                const int r = side ? rn : row;
                const ModeK<T> k = mode_k<T>(a.ky, sn[side], kxr[side], iky);
                a.aphi[(long long)r * g.nky + iky] = (r == 0 && iky == 0) ? czero<T>() : cdivr(base[m], k.k2);
            }
            if (tl == 0 && row != 0) {
                const ModeK<T> k = mode_k<T>(a.ky, sn[1], kxr[1], 0);
                a.aphi[(long long)rn * g.nky] = cdivr(base_n0, k.k2);
            }
        }
        TeamFFT<T, NY, P, S::RMAX, +1>::run(v, lds, a.twy, tid, sync);
        cplx<T>* Gf = a.G + (long long)oi * g.I;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int y = tl + m * TT;
            Gf[((long long)(y >> 2) * g.nkxh + gp) * 4 + (y & 3)] = v[m];
        }
    }
}

// P1 with one (row pair, field group) per workgroup: the row pair of omega
// (groups 0: vx, vy and 1: d/dx, d/dy omega) or rho (group 2) is read once
// through the remap view and both outputs of the group are transformed from
// it.  Row 0 (kx = 0) has three transformed outputs, vx, d/dy omega and d/dy
// rho (the others carry the factor kx = 0 and are written as zeros): its
// omega workgroup transforms vx and d/dy omega, its rho workgroup d/dy rho and
// writes the zeros.  The grid is 3 (nkxh - 1) + 2 workgroups, 2048 at 2048^2:
// all resident at once at eight per CU (3 nkxh = 2049 left one workgroup to
// start after the first ones finished).  Linear mode: group 0 only, nkxh
// workgroups.
template <typename T, int NY, bool BATCH>
__global__ void __launch_bounds__((TeamShape<T, NY>::TT), (waves_per_eu<T, NY, TeamShape<T, NY>::TT, 1>()))
k_p1_pair(StepArgs<T> a1, const StepArgs<T>* __restrict__ args) {
    const StepArgs<T>& a = step_args_of<BATCH>(a1, args);
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    constexpr int NKY = NY / 3 + 1;
    __shared__ cplx<T> sm[L];
    const Geo& g = a.g;
    const int tid = threadIdx.x;
    const int ngrp = a.ph.linear ? 1 : 3;
    const int q = xcd_item(blockIdx.x, gridDim.x);
    const int nreg = ngrp * (g.nkxh - 1);                 // items of rows 1 .. nkxh-1
    const bool r0 = !a.ph.linear && q >= nreg;            // row 0: omega (q = nreg) or rho (nreg + 1)
    const int pos = r0 ? g.nkxh - 1 : q / ngrp;
    const int grp = r0 ? 2 * (q - nreg) : q - pos * ngrp;
    const unsigned omask = r0 ? (grp == 0 ? 0x9u : 0x20u) : (3u << (2 * grp));
    const int row = p1_row_of(pos, g.nkxh);
    const int rn = row == 0 ? 0 : g.nkx - row;
    const StepCtl* c = a.ctl_in;
    const int s1 = slot_a1(c->istep);
    const cplx<T>* src = grp == 2 ? a.a_rho[s1] : a.a_omg[s1];
    const int* jmp = (c->slot_epoch_a[s1] != c->epoch) ? a.jump + (c->epoch & 1) * g.nkx : nullptr;
    const double sn[2] = {a.s_in[row], a.s_in[rn]};
    const double kxr[2] = {a.kx[row], a.kx[rn]};
    cplx<T> base[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = tid + m * TT;
        const bool p0 = e < NKY;
        const bool keep = p0 || e > NY - NKY;
        const int iky = keep ? (p0 ? e : NY - e) : 0;
        const cplx<T> F = view_load(src, p0 ? row : rn, iky, g, jmp);
        base[m] = keep ? F : czero<T>();
    }
    const cplx<T> base_n0 = view_load(src, rn, 0, g, jmp);
    p1_outputs<T, NY>(a, row, base, base_n0, sn, kxr, omask, a.store_phi, sm, tid);
    if (r0 && grp == 2) {
        // outputs of row 0 that carry a factor kx = 0: vy, d/dx omega, d/dx rho
        for (int idx = tid; idx < 3 * NY; idx += TT) {
            const int z = idx / NY, y = idx - z * NY;
            const int o = z == 0 ? 1 : (z == 1 ? 2 : 4);
            a.G[(long long)o * g.I + ((long long)(y >> 2) * g.nkxh + pos) * 4 + (y & 3)] = czero<T>();
        }
    }
}

// max(m, |x|) in the precision of its operands (NaN x leaves m).
__device__ __forceinline__ float maxabs(float m, float x) { return fmaxf(m, fabsf(x)); }
__device__ __forceinline__ double maxabs(double m, double x) { return fmax(m, fabs(x)); }

// ---------------------------------------------------------------------------
// P2
// ---------------------------------------------------------------------------
// Workgroups of P2: KT = p2_teams columns each (a quarter, half or all of a
// G tile of kTile columns); the workgroups of one tile run on one XCD so the
// tile's lines are fetched once into its L2.
template <typename T, int NX, bool BATCH, bool THREE>
__global__ void __launch_bounds__((p2_teams<T, NX>() * Team<NX, p2_points<T, NX>()>::TT),
                                  (waves_per_eu<T, NX, Team<NX, p2_points<T, NX>()>::TT, p2_teams<T, NX>()>()))
k_p2(StepArgs<T> a1, const StepArgs<T>* __restrict__ args) {
    const StepArgs<T>& a = step_args_of<BATCH>(a1, args);
    using S = Team<NX, p2_points<T, NX>()>;
    constexpr int KT = p2_teams<T, NX>();
    constexpr int SUB = kTile / KT;                    // workgroups per G tile
    constexpr int P = S::P, TT = S::TT, L = lds_len<NX>();
    constexpr int NT = KT * TT;
    constexpr int NKMAX = NX / 3 + 1;                  // >= nkxh
    constexpr int NQ = (NKMAX + TT - 1) / TT;          // kx per thread in the split
    constexpr int NS = (KT * NKMAX + NT - 1) / NT;     // staged elements per thread per field
    constexpr int KP = p2_stage_stride<NX>();
    static_assert(kTile % KT == 0, "P2 workgroups split a G tile evenly");
    __shared__ cplx<T> sm[KT * L];
    __shared__ double red[32];
    const Geo& g = a.g;
    const int team = threadIdx.x / TT, tid = threadIdx.x - team * TT;
    const int q = xcd_item(blockIdx.x, gridDim.x);
    const int tile = q / SUB;
    const int cbase = (q - tile * SUB) * KT;           // first column of this workgroup in the tile
    const int y = tile * kTile + cbase + team;
    constexpr int nk = NKMAX;   // == nkxh for the supported (power-of-two) lengths
    cplx<T>* lds = sm + team * L;
    const TeamSync<S::TT> sync{};

    // CFL maxima (add_v0 + pl_max_search): max|vx + sigma*yy[y]|, max|vy|, one
    // partial pair per workgroup (P3 reduces the array; no atomics)
    // (in the working precision, as add_v0 forms vx + sigma*yy in cureal; the
    // maxima are exact and converted to FP64 once per workgroup)
    const T sy = (T)mul_rn(a.ph.sigma, (double)y * g.dy);
    T m0 = 0, m1 = 0;

    // Inverse x-transforms of G field pairs FA + i FB for this team's column.
    // Four-pass form: (vx, dw/dx), (vy, dw/dy), (vx, dr/dx), (vy, dr/dy), so
    // each velocity meets the derivative it multiplies and only the two
    // brackets stay live between transforms.  Three-pass form (THREE): (vx,
    // vy), (dw/dx, dw/dy), (dr/dx, dr/dy).  Linear: (vx, vy) for the CFL maxima
    // only.
    // The fields of G: 0 vx = X(i kys phi), 1 vy = X(-i kx phi), 2 X(i kx w),
    // 3 X(i kys w), 4 X(i kx r), 5 X(i kys r).  The loop is not unrolled and
    // its lane index is made opaque per pass, so addresses are recomputed
    // rather than held in registers across transforms.
    // THREE: the pairs (vx, vy), (dw/dx, dw/dy), (dr/dx, dr/dy), one transform
    // fewer, with vx, vy and B_omega live across the last one.
    T bw[P], br[P];
    T vxs[THREE ? P : 1];
    const int npass = a.ph.linear ? 1 : (THREE ? 3 : 4);
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
        int lane = threadIdx.x;
        asm volatile("" : "+v"(lane));
        const int tl = (lane - team * TT) & (TT - 1);
        const int fa = a.ph.linear ? 0 : (THREE ? 2 * pass : (pass & 1));
        const int fb = a.ph.linear ? 1 : (THREE ? 2 * pass + 1 : 2 + pass);
        const cplx<T>* GA = a.G + (long long)fa * g.I + (long long)tile * nk * 4;
        const cplx<T>* GB = a.G + (long long)fb * g.I + (long long)tile * nk * 4;
        // stage the two tiles as [field][c][kx], all loads first
        cplx<T> ra[NS], rb[NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) {   // element idx: position idx / KT, column cbase + idx % KT
            const int idx = lane + m * NT;
            const int gi = idx < nk * KT ? (idx / KT) * kTile + cbase + idx % KT : 0;
            ra[m] = GA[gi];
            rb[m] = GB[gi];
        }
        teams_join<TT>();
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            const int idx = lane + m * NT;
            if (idx < nk * KT) {
                sm[(idx % KT) * KP + idx / KT] = ra[m];
                sm[(KT + idx % KT) * KP + idx / KT] = rb[m];
            }
        }
        __syncthreads();
        // Z(kx) = A + iB for kx >= 0, conj(A) + i conj(B) at nx - kx,
        // branch-free; real parts only at kx = 0
        const cplx<T>* ca = sm + team * KP;
        const cplx<T>* cb = sm + (KT + team) * KP;
        cplx<T> v[P];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int e = tl + m * TT;
            const bool pos = e < nk;
            const bool keep = pos || e > NX - nk;
            const int kk = keep ? (pos ? e : NX - e) : 0;
            const int gp = g_pos_of(kk, nk);   // G position of row kk
            const cplx<T> A = ca[gp], B = cb[gp];
            const cplx<T> z = pos ? cplx<T>{A.x - B.y, A.y + B.x} : cplx<T>{A.x + B.y, B.x - A.y};
            v[m] = keep ? z : czero<T>();
        }
        if (tl == 0) v[0] = {ca[nk - 1].x, cb[nk - 1].x};   // kx = 0: real parts only
        __syncthreads();
        TeamFFT<T, NX, P, S::RMAX, +1>::run(v, lds, a.twx, tl, sync);

        if (a.ph.linear) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                m0 = maxabs(m0, v[m].x + sy);
                m1 = maxabs(m1, v[m].y);
            }
        } else if (THREE) {
            // bw holds vx until B_omega exists, br holds vy until B_rho replaces it
            if (pass == 0) {
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    m0 = maxabs(m0, v[m].x + sy);
                    m1 = maxabs(m1, v[m].y);
                    vxs[m] = v[m].x;
                    br[m] = v[m].y;
                }
            } else if (pass == 1) {
#pragma unroll
                for (int m = 0; m < P; ++m) bw[m] = vxs[m] * v[m].x + br[m] * v[m].y;
            } else {
#pragma unroll
                for (int m = 0; m < P; ++m) br[m] = vxs[m] * v[m].x + br[m] * v[m].y;
            }
        } else if (pass == 0) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                m0 = maxabs(m0, v[m].x + sy);
                bw[m] = v[m].x * v[m].y;
            }
        } else if (pass == 1) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                m1 = maxabs(m1, v[m].x);
                bw[m] += v[m].x * v[m].y;
            }
        } else if (pass == 2) {
#pragma unroll
            for (int m = 0; m < P; ++m) br[m] = v[m].x * v[m].y;
        } else {
#pragma unroll
            for (int m = 0; m < P; ++m) br[m] += v[m].x * v[m].y;
        }
    }

    T c0 = wave_max(m0), c1 = wave_max(m1);
    {
        const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
        if ((threadIdx.x & 63) == 0) { red[w] = c0; red[16 + w] = c1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < nw; ++i) { c0 = fmax(c0, (T)red[i]); c1 = fmax(c1, (T)red[16 + i]); }
            reinterpret_cast<cplx<T>*>(a.cflp)[q] = cplx<T>{c0, c1};
        }
    }
    if (a.ph.linear) return;

    // the negated brackets (negative) as z = -B_omega - i B_rho
    cplx<T> v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = {-bw[m], -br[m]};

    // x-forward of z = -B_omega - i B_rho, then split into the two spectra
    TeamFFT<T, NX, P, S::RMAX, -1>::run(v, lds, a.twx, tid, sync);
#pragma unroll
    for (int m = 0; m < P; ++m) lds[tid + m * TT] = v[m];
    __syncthreads();
    cplx<T> hw[NQ], hr[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int kx = tid + i * TT;
        if (kx < nk) {
            const cplx<T> Zk = lds[kx], Zm = lds[(NX - kx) & (NX - 1)];
            hw[i] = {(Zk.x + Zm.x) * T(0.5), (Zk.y - Zm.y) * T(0.5)};
            hr[i] = {(Zk.y + Zm.y) * T(0.5), (Zm.x - Zk.x) * T(0.5)};
        }
    }
    __syncthreads();
    // stage as [field][c][kx], then write H as [t][kx][c] tiles
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int kx = tid + i * TT;
        if (kx < nk) {
            sm[team * KP + kx] = hw[i];
            sm[(KT + team) * KP + kx] = hr[i];
        }
    }
    __syncthreads();
    cplx<T>* Hw = a.H + (long long)tile * nk * 4;
    cplx<T>* Hr = Hw + g.I;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int idx = threadIdx.x + m * NT;
        if (idx < nk * KT) {
            const int gi = (idx / KT) * kTile + cbase + idx % KT;
            Hw[gi] = sm[(idx % KT) * KP + idx / KT];
            Hr[gi] = sm[(KT + idx % KT) * KP + idx / KT];
        }
    }
}

// ---------------------------------------------------------------------------
// P3
// ---------------------------------------------------------------------------
// Explicit step + implicit diffusion for one mode (time_integral.cu:621-695,
// 747-767), in the reference's operation order.  dt12 = dt / 12 and dt9 =
// dt / 9 come from the caller, formed once per workgroup.
template <typename T>
__device__ __forceinline__ cplx<T> ab_bdf_update(int ist, bool plain, double dt, double dt12, double dt9,
                                                 double diff, double k2, cplx<T> a1, cplx<T> a2, cplx<T> d0,
                                                 cplx<T> d1, cplx<T> d2) {
    cplx<T> r;
    if (ist == 0) {                                            // eulerf
        const T c = (T)dt;
        r = {a1.x + c * d0.x, a1.y + c * d0.y};
    } else if (ist == 1) {
        if (plain) {                                           // ab2
            const T c = (T)dt;
            const cplx<T> w = {T(1.5) * d0.x - T(0.5) * d1.x, T(1.5) * d0.y - T(0.5) * d1.y};
            r = {a1.x + c * w.x, a1.y + c * w.y};
        } else {                                               // ab2, BDF2 form
            const T c = (T)((2.0 / 3.0) * dt);
            const cplx<T> u = {(T(4) * a1.x - a2.x) / T(3), (T(4) * a1.y - a2.y) / T(3)};
            const cplx<T> w = {T(2) * d0.x - d1.x, T(2) * d0.y - d1.y};
            r = {u.x + c * w.x, u.y + c * w.y};
        }
    } else {
        if (plain) {                                           // ab3
            const T c = (T)dt12;
            const cplx<T> w = {(T(23) * d0.x - T(16) * d1.x) + T(5) * d2.x,
                               (T(23) * d0.y - T(16) * d1.y) + T(5) * d2.y};
            r = {a1.x + c * w.x, a1.y + c * w.y};
        } else {                                               // ab3, BDF2 form
            const T c = (T)dt9;
            const cplx<T> u = {(T(4) * a1.x - a2.x) / T(3), (T(4) * a1.y - a2.y) / T(3)};
            const cplx<T> w = {(T(16) * d0.x - T(14) * d1.x) + T(4) * d2.x,
                               (T(16) * d0.y - T(14) * d1.y) + T(4) * d2.y};
            r = {u.x + c * w.x, u.y + c * w.y};
        }
    }
    if (!plain) {                                              // eulerb / bdf2
        const double cd = ist == 0 ? dt : mul_rn(2.0 / 3.0, dt);
        const T den = (T)add_rn(1.0, mul_rn(mul_rn(cd, diff), k2));
        r = cdivr(r, den);
    }
    return r;
}

// FP32 contexts: the state update (AB/BDF2 combination and implicit
// diffusion) is evaluated in FP64 from the FP32 operands and rounded once,
// stochastically, so each step adds one unbiased rounding of the state instead
// of about five round-to-nearest ones whose sub-ulp bias accumulates over
// thousands of small steps; the tendency, which enters multiplied by the step,
// stays FP32.
constexpr bool kWideUpdate = true;

// 32-bit integer hash (the "lowbias32" finaliser of Chris Wellons' hash
// prospector: xor-shift / multiply rounds).
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// FP64 -> FP32 with stochastic rounding driven by a deterministic key: x
// rounds away from zero with probability equal to its distance from the float
// below it in magnitude over the gap (Connolly, Higham and Mary, "Stochastic
// rounding and its probabilistic backward error analysis", 2021), done by
// adding a uniform 29-bit integer to the 29 mantissa bits of the double that
// a float drops and clearing them, after which the conversion is exact.  The
// expected result is x, so the sub-ulp changes of a slowly evolving FP32
// state accumulate instead of being rounded away every step.  Infinities and
// NaNs pass unchanged; results in the float subnormal range (no state value
// here comes near it) round to nearest after the truncation.
__device__ __forceinline__ float round_stochastic(double x, uint32_t key) {
    const unsigned long long noise = isfinite(x) ? (hash32(key) >> 3) : 0u;
    const unsigned long long b = ((unsigned long long)__double_as_longlong(x) + noise) & ~0x1fffffffull;
    return (float)__longlong_as_double((long long)b);
}

// The same update for an FP32 state evaluated in FP64 and rounded once, with
// stochastic rounding keyed by (step, field, mode): the divisions by 3 and by
// the diffusion denominator become one reciprocal and products (their FP64
// rounding is far below the final FP32 rounding).
__device__ __forceinline__ cplx<float> ab_bdf_update_wide(int ist, bool plain, double dt, double dt12, double dt9,
                                                         double diff, double k2,
                                                         cplx<float> a1f, cplx<float> a2f, cplx<float> d0f,
                                                         cplx<float> d1f, cplx<float> d2f, uint32_t key) {
    const cplx<double> a1 = {a1f.x, a1f.y}, a2 = {a2f.x, a2f.y};
    const cplx<double> d0 = {d0f.x, d0f.y}, d1 = {d1f.x, d1f.y}, d2 = {d2f.x, d2f.y};
    constexpr double third = 1.0 / 3.0;
    cplx<double> r;
    if (ist == 0) {
        r = {a1.x + dt * d0.x, a1.y + dt * d0.y};
    } else if (ist == 1) {
        if (plain) {
            r = {a1.x + dt * (1.5 * d0.x - 0.5 * d1.x), a1.y + dt * (1.5 * d0.y - 0.5 * d1.y)};
        } else {
            const double c = (2.0 / 3.0) * dt;
            r = {(4.0 * a1.x - a2.x) * third + c * (2.0 * d0.x - d1.x),
                 (4.0 * a1.y - a2.y) * third + c * (2.0 * d0.y - d1.y)};
        }
    } else {
        if (plain) {
            const double c = dt12;
            r = {a1.x + c * ((23.0 * d0.x - 16.0 * d1.x) + 5.0 * d2.x), a1.y + c * ((23.0 * d0.y - 16.0 * d1.y) + 5.0 * d2.y)};
        } else {
            const double c = dt9;
            r = {(4.0 * a1.x - a2.x) * third + c * ((16.0 * d0.x - 14.0 * d1.x) + 4.0 * d2.x),
                 (4.0 * a1.y - a2.y) * third + c * ((16.0 * d0.y - 14.0 * d1.y) + 4.0 * d2.y)};
        }
    }
    if (!plain) {
        const double cd = ist == 0 ? dt : (2.0 / 3.0) * dt;
        const double den = 1.0 + cd * diff * k2;
        // 1/den: FP32 reciprocal refined by one Newton step (relative error ~1e-14)
        double inv = (double)__builtin_amdgcn_rcpf((float)den);
        inv = inv * (2.0 - den * inv);
        r = {r.x * inv, r.y * inv};
    }
    return {round_stochastic(r.x, key), round_stochastic(r.y, key * 0x2c1b3c6du + 0x297a2d39u)};
}

// Bookkeeping of step n done by workgroup 0 of P3: the shift advance and remap
// trigger of update_shear (shear.cu:211-294) in FP64, and the next control block.
template <typename T>
__device__ void p3_bookkeeping(const StepArgs<T>& a, double dt, int h, int which, double cvx, double cvy) {
    const Geo& g = a.g;
    const StepCtl* c = a.ctl_in;
    const double sdt1 = g.nkx > 1 ? mul_rn(mul_rn(a.ph.sigma, a.kx[1]), dt) : 0.0;
    const double s1n = g.nkx > 1 ? sub_rn(a.s_in[1], sdt1) : 0.0;
    const int trig = (g.nkx > 1 && fabs(sub_rn(s1n, -a.kx[1])) < 1e-10) ? 1 : 0;
    const int newE = c->epoch + trig;
    int* jn = a.jump + (newE & 1) * g.nkx;
    for (int i = threadIdx.x; i < g.nkx; i += blockDim.x) {
        double sn = a.s_in[i];
        if (i > 0) sn = sub_rn(a.s_in[i], mul_rn(mul_rn(a.ph.sigma, a.kx[i]), dt));
        if (trig) {
            jn[i] = i == 0 ? 0 : (int)round(sn / a.kx[1]);
            if (i > 0) sn = 0.0;
        }
        a.s_out[i] = sn;
    }
    if (threadIdx.x == 0) {
        // field by field into ctl_out (no private copy with dynamically indexed arrays)
        const long long n = c->istep;
        StepCtl* o = a.ctl_out;
        o->time = add_rn(c->time, dt);
        o->delt = dt;
        o->istep = n + 1;
        o->epoch = newE;
        const int s0 = slot_a0(n), s1 = slot_a1(n);
        o->error = c->error | ((trig && (c->slot_epoch_a[s1] != c->epoch || c->slot_epoch_d[s1] != c->epoch)) ? 1 : 0);
        for (int s = 0; s < 3; ++s) {
            o->slot_epoch_a[s] = s == s0 ? c->epoch : c->slot_epoch_a[s];
            o->slot_epoch_d[s] = s == s0 ? c->epoch : c->slot_epoch_d[s];
        }
        // aphi's frame: this step's P1 (store_phi) or, fused, the next step's (store_phi_next)
        o->phi_epoch = a.store_phi_next ? newE : (a.store_phi ? c->epoch : c->phi_epoch);
        o->remap_count = c->remap_count + trig;
        o->last_remap_step = trig ? n : c->last_remap_step;
        o->halvings = c->halvings + h;
        o->log_count = c->log_count;
        if (h > 0 && a.hlog != nullptr && c->log_count < a.hlog_cap) {
            const double value = which == 0 ? cvx
                               : which == 1 ? cvy
                               : which == 2 ? a.ph.g / a.ph.rho0
                               : which == 3 ? a.ph.rho0_prime : 0.0;
            HalvingRec rec;
            rec.istep = n;
            rec.time = c->time;
            rec.value = value;
            rec.delt = dt;
            rec.which = which;
            rec.pad_ = h;
            double dtc;
            int wc;
            cfl_delt(c, g, a.ph, cvx, cvy, &dtc, &wc, rec.counts);
            rec.values[0] = cvx;
            rec.values[1] = cvy;
            rec.values[2] = a.ph.g / a.ph.rho0;
            rec.values[3] = a.ph.rho0_prime;
            rec.delt_before = c->delt;
            a.hlog[c->log_count] = rec;
            o->log_count = c->log_count + 1;
        }
    }
}

// Workgroups of P3: (group of KT row pairs, field), KT = 1 or kTile.  One row
// pair per workgroup (KT = 1) gives eight small workgroups per CU instead of
// two large ones, so their load, transform and update phases overlap.
//
// FUSE (KT = 1): the workgroup goes on to the next step's P1 for its row pair
// and field -- the "fused spectral kernel" of one step boundary.  The update
// leaves the new rows in LDS (in place of the bracket values it consumed,
// each position read and then written by the same thread; the two ky = 0
// values in their own slots), the next shift and remap trigger follow from
// the same FP64 operations as the bookkeeping, and the outputs of this field
// (omega: vx, vy, d/dx, d/dy; rho: d/dx, d/dy) are transformed and written to
// G through the remap view, as k_p1_row would after reading the rows back.
// SPLIT: two workgroups per (row pair, field), one per row of the pair (side
// 0: kx, side 1: -kx); each transforms the H row (the transform gives both
// rows) and updates its own row only.  Twice the workgroups for grids whose
// row pairs alone do not fill the chip; the same arithmetic per mode.
template <typename T, int NY, bool BATCH, int KT, bool FUSE, bool NARROW = false, bool SPLIT = false>
__global__ void __launch_bounds__((KT * TeamShape<T, NY>::TT), (waves_per_eu<T, NY, TeamShape<T, NY>::TT, KT>()))
k_p3(StepArgs<T> a1, const StepArgs<T>* __restrict__ args) {
    const StepArgs<T>& a = step_args_of<BATCH>(a1, args);
    using S = TeamShape<T, NY>;
    constexpr int P = S::P, TT = S::TT, L = lds_len<NY>();
    constexpr int NT = KT * TT;
    constexpr int LP = stage_stride<NY>();
    constexpr int NKY = NY / 3 + 1;                 // == nky for the supported (power-of-two) lengths
    constexpr int J = (NKY + TT - 1) / TT;          // row points per thread
    // points whose loads are in flight together: three in the plain FP32 pair
    // kernel, two where the batched argument block, the four-row tile or the
    // fused tail need the registers
    constexpr int JB = (sizeof(T) == 4 && !BATCH && KT == 1 && !FUSE) ? 3 : 2;
    static_assert(KT == 1 || KT == kTile, "P3 workgroups hold one row pair or one H tile");
    static_assert(!FUSE || KT == 1, "the fused next-step P1 works on one row pair");
    static_assert(!SPLIT || (KT == 1 && !FUSE), "split workgroups hold one row of one pair");
    __shared__ cplx<T> sm[KT * L];
    __shared__ double red[32];
    __shared__ cplx<T> slot[2];   // FUSE: the new ky = 0 values of the row and of its mirror
    const Geo& g = a.g;
    const Phys& ph = a.ph;
    const int team = KT == 1 ? 0 : (int)threadIdx.x / TT, tid = threadIdx.x - team * TT;
    // the omega and rho workgroups of a row group read each other's rows: keep them on one XCD
    const int qq = xcd_item(blockIdx.x, gridDim.x);
    const int myside = SPLIT ? (qq & 1) : 0;   // SPLIT: the row of the pair this workgroup updates
    const int q = SPLIT ? (qq >> 1) : qq;
    // FUSE: three workgroups per row pair kx >= 1 -- omega with the next step's
    // vx, vy; omega again with d/dx, d/dy omega (the same update, not stored);
    // rho with d/dx, d/dy rho -- and two for row 0 (omega: vx, d/dy omega;
    // rho: d/dy rho and the zero outputs), 3 (nkxh - 1) + 2 in all
    const int nreg3 = 3 * (g.nkxh - 1);
    const int fk = FUSE ? (q < nreg3 ? q % 3 : 2 * (q - nreg3)) : 0;
    const int tile = FUSE ? (q < nreg3 ? q / 3 + 1 : 0) : q >> 1;
    const int f = FUSE ? (fk == 2 ? 1 : 0) : (q & 1);   // 0: omega, 1: rho
    const bool writer = !FUSE || fk != 1;               // stores the new state and tendency
    const int row = tile * KT + team;
    const bool act = row < g.nkxh;
    const int rn = row == 0 ? 0 : g.nkx - row;
    const int nrow_tile = min(KT, g.nkxh - tile * KT);
    const StepCtl* c = a.ctl_in;
    const long long n = c->istep;
    const uint32_t step_key = hash32((uint32_t)n * 2u + (uint32_t)f);   // stochastic-rounding stream of (step, field)
    cplx<T>* lds = sm + team * L;
    const TeamSync<S::TT> sync{};

    const int s0 = slot_a0(n), s1 = slot_a1(n), s2 = slot_a2(n);
    const int* jbase = a.jump + (c->epoch & 1) * g.nkx;
    const int* ja1 = c->slot_epoch_a[s1] != c->epoch ? jbase : nullptr;
    const int* ja2 = c->slot_epoch_a[s2] != c->epoch ? jbase : nullptr;
    const int* jd1 = c->slot_epoch_d[s1] != c->epoch ? jbase : nullptr;
    const int* jd2 = c->slot_epoch_d[s2] != c->epoch ? jbase : nullptr;
    const int ist = n == 0 ? 0 : (n == 1 ? 1 : 2);
    const T scale = (T)g.scale;
    double dt = 0.0, dt12 = 0.0, dt9 = 0.0, cvx = 0.0, cvy = 0.0;
    int which = -1;

    cplx<T>* A0 = f ? a.a_rho[s0] : a.a_omg[s0];
    const cplx<T>* A1 = f ? a.a_rho[s1] : a.a_omg[s1];
    const cplx<T>* A2 = f ? a.a_rho[s2] : a.a_omg[s2];
    cplx<T>* D0 = f ? a.d_rho[s0] : a.d_omg[s0];
    const cplx<T>* D1 = f ? a.d_rho[s1] : a.d_omg[s1];
    const cplx<T>* D2 = f ? a.d_rho[s2] : a.d_omg[s2];
    const cplx<T>* X1 = f ? a.a_omg[s1] : a.a_rho[s1];   // the other field at step n
    const double diff = f ? ph.kappa : ph.nu;
    const bool plain = diff < ph.eps;

    // The update streams its operands in batches of JB points of one row (side
    // 0: kx, side 1: -kx); the loads of batch b+1 are issued before batch b is
    // computed, and those of batch 0 before the H transform, so the transform
    // runs while the first operands arrive.
    constexpr int NBS = (J + JB - 1) / JB;   // batches per side
    constexpr int NB = 2 * NBS;
    struct Batch {
        cplx<T> x1[JB], a1[JB], a2[JB], d1[JB], d2[JB];
        double ky[JB];
    };
    // kx and the shift of the two rows, loaded before anything is in flight
    const int rc = act ? row : 0;
    const double kx_r[2] = {a.kx[rc], a.kx[rc == 0 ? 0 : g.nkx - rc]};
    const double s_r[2] = {a.s_in[rc], a.s_in[rc == 0 ? 0 : g.nkx - rc]};
    // Mode (r, ik) of a ring buffer: through the remap view only in the steps
    // right after a remap (VIEW), a plain load otherwise.
    auto ld = [&](auto view, const cplx<T>* buf, int r, int ik, const int* jmp) -> cplx<T> {
        if constexpr (decltype(view)::value) return view_load(buf, r, ik, g, jmp);
        else return buf[(long long)r * NKY + ik];
    };
    auto load_batch = [&](auto view, auto bc, Batch& B) {
        constexpr int b = decltype(bc)::value, j0 = (b % NBS) * JB;
        const int side = SPLIT ? myside : b / NBS;
        if (!act || (side == 1 && row == 0)) return;
        const int r = side ? rn : row;
#pragma unroll
        for (int u = 0; u < JB; ++u) {
            const int iky = tid + (j0 + u) * TT;
            const int ik = (j0 + u < J && iky < NKY) ? iky : 0;   // in bounds for idle lanes
            // unconditional: the first steps load history slots they then ignore
            B.x1[u] = ld(view, X1, r, ik, ja1);
            B.a1[u] = ld(view, A1, r, ik, ja1);
            B.a2[u] = ld(view, A2, r, ik, ja2);
            B.d1[u] = ld(view, D1, r, ik, jd1);
            B.d2[u] = ld(view, D2, r, ik, jd2);
            B.ky[u] = a.ky[ik];
        }
    };
    auto update_batch = [&](auto bc, const Batch& B) {
        constexpr int b = decltype(bc)::value, j0 = (b % NBS) * JB;
        const int side = SPLIT ? myside : b / NBS;
        if (!act || (side == 1 && row == 0)) return;
        const int r = side ? rn : row;
        const double kxv = side ? kx_r[1] : kx_r[0];   // no dynamic index into the two-element arrays
        const double sr = side ? s_r[1] : s_r[0];
        // add_ddxrho: g/rho0 * kx;  add_ddxphi: rho0' * kx
        const T cc = (T)mul_rn(f == 0 ? ph.g / ph.rho0 : ph.rho0_prime, kxv);
#pragma unroll
        for (int u = 0; u < JB; ++u) {
            const int iky = tid + (j0 + u) * TT;
            if (j0 + u >= J || iky >= NKY) continue;
            cplx<T> nh = czero<T>();
            if (!ph.linear) {
                const cplx<T> F = side ? cconj(lds[(NY - iky) & (NY - 1)]) : lds[iky];
                nh = {F.x * scale, F.y * scale};
            }
            const double k2 = k2_of(kxv, add_rn(B.ky[u], sr));
            cplx<T> d0, an;
            if (f == 0) {
                d0 = {nh.x + cc * B.x1[u].y, nh.y - cc * B.x1[u].x};
            } else {   // phi = omega/kperp2 (zero at the origin)
                const cplx<T> phi = (r == 0 && iky == 0) ? czero<T>() : cdivr(B.x1[u], (T)k2);
                d0 = {nh.x - cc * phi.y, nh.y + cc * phi.x};
            }
            if constexpr (sizeof(T) == 4 && kWideUpdate && !NARROW) {
                // FP32 state: the AB/BDF2 combination and the implicit diffusion in FP64, rounded once
                const uint32_t key = step_key + (uint32_t)(r * NKY + iky) * 0x9e3779b1u;
                an = ab_bdf_update_wide(ist, plain, dt, dt12, dt9, diff, k2, B.a1[u], B.a2[u], d0, B.d1[u], B.d2[u],
                                        key);
            } else {
                an = ab_bdf_update<T>(ist, plain, dt, dt12, dt9, diff, k2, B.a1[u], B.a2[u], d0, B.d1[u], B.d2[u]);
            }
            const long long at = (long long)r * NKY + iky;
            if (writer) {
                A0[at] = an;
                D0[at] = d0;
            }
            if constexpr (FUSE) {
                if (iky == 0) slot[side] = an;
                else lds[side ? NY - iky : iky] = an;   // the position this thread just read
            }
        }
    };

    auto run = [&](auto view) {
        Batch bt[2];

        // the step size after the CFL halvings, and (workgroup 0) the bookkeeping
        cfl_maxima<T, NT>(a.cflp, a.ncflp, red, &cvx, &cvy);
        const int h = cfl_delt(c, g, ph, cvx, cvy, &dt, &which);
        dt12 = dt / 12.0;   // the AB3 weights' factors, once per workgroup
        dt9 = dt / 9.0;
        if (blockIdx.x == 0) p3_bookkeeping(a, dt, h, which, cvx, cvy);

        if (!ph.linear) {
            // twiddles and the H tile of this field, then [team][y] in LDS
            using FFT = TeamFFT<T, NY, P, S::RMAX, -1>;
            typename FFT::Twiddles ts;
            ts.load(a.twy, tid);
            const cplx<T>* Hf = a.H + (long long)f * g.I;
            cplx<T> v[P];
            if constexpr (KT == 1) {
                // this row of the [t][kx][c] tiles: runs of 4 (32 bytes) per t
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    const int y = tid + m * TT;
                    v[m] = Hf[((long long)(y >> 2) * g.nkxh + row) * 4 + (y & 3)];
                }
            } else {
                cplx<T> hv[P];
#pragma unroll
                for (int m = 0; m < P; ++m) {   // NY*4 tile elements over NT threads
                    const int idx = threadIdx.x + m * NT;
                    const int t = idx >> 4, w = idx & 15;
                    hv[m] = (w >> 2) < nrow_tile ? Hf[((long long)t * g.nkxh + tile * KT) * 4 + w] : czero<T>();
                }
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    const int idx = threadIdx.x + m * NT;
                    const int t = idx >> 4, w = idx & 15;
                    sm[(w >> 2) * LP + 4 * t + (w & 3)] = hv[m];
                }
                __syncthreads();
#pragma unroll
                for (int m = 0; m < P; ++m) v[m] = sm[team * LP + tid + m * TT];
                __syncthreads();
            }
            // the first operands travel during the transform (loads complete in
            // order, so they are issued after the H tile has been consumed)
            load_batch(view, std::integral_constant<int, 0>{}, bt[0]);
            FFT::run(v, lds, ts, tid, sync);
#pragma unroll
            for (int m = 0; m < P; ++m) lds[tid + m * TT] = v[m];
            __syncthreads();
        } else {
            load_batch(view, std::integral_constant<int, 0>{}, bt[0]);
        }

        constexpr int NBT = SPLIT ? NBS : NB;   // the batches of this workgroup's row(s)
        static_for<0, NBT>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            if constexpr (b + 1 < NBT) load_batch(view, std::integral_constant<int, b + 1>{}, bt[(b + 1) & 1]);
            asm volatile("" ::: "memory");   // keep those loads ahead of this batch's stores
            update_batch(bc, bt[b & 1]);
        });
    };
    if (ja1 != nullptr || ja2 != nullptr || jd1 != nullptr || jd2 != nullptr) run(std::true_type{});
    else run(std::false_type{});

    if constexpr (FUSE) {
        if (!act || (ph.linear && f == 1)) return;
        if (ph.linear && fk == 1) return;   // no d/dx, d/dy outputs in the linear mode
        __syncthreads();   // the new rows are in LDS
        // shift and remap of this step for the two rows (p3_bookkeeping's operations)
        const double sdt1 = g.nkx > 1 ? mul_rn(mul_rn(ph.sigma, a.kx[1]), dt) : 0.0;
        const double s1n = g.nkx > 1 ? sub_rn(a.s_in[1], sdt1) : 0.0;
        const bool trig = g.nkx > 1 && fabs(sub_rn(s1n, -a.kx[1])) < 1e-10;
        double sn[2];
        int jv[2];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int r = side ? rn : row;
            double v = s_r[side];
            if (r > 0) v = sub_rn(s_r[side], mul_rn(mul_rn(ph.sigma, kx_r[side]), dt));
            jv[side] = (trig && r > 0) ? (int)round(v / a.kx[1]) : 0;
            sn[side] = (trig && r > 0) ? 0.0 : v;
        }
        // the new rows in the frame of step n+1 (view_load's remap on the LDS rows)
        auto rowval = [&](int side, int ky) -> cplx<T> {
            if (row == 0) side = 0;   // row 0 is its own mirror
            if (ky == 0) return slot[side];
            return side ? lds[NY - ky] : lds[ky];
        };
        auto viewval = [&](int side, int ky) -> cplx<T> {
            if (!trig) return rowval(side, ky);
            const int j = jv[side];
            if (j <= 0) return ky - j < NKY ? rowval(side, ky - j) : czero<T>();
            if (ky - j > 0) return rowval(side, ky - j);
            return j - ky < NKY ? cconj(rowval(1 - side, j - ky)) : czero<T>();
        };
        // the spectrum of the row pair before multipliers: base[m] for e = tid + m*TT
        cplx<T> base[P];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const int e = tid + m * TT;
            const bool pos = e < NKY;
            const bool keep = pos || e > NY - NKY;
            base[m] = keep ? viewval(pos ? 0 : 1, pos ? e : NY - e) : czero<T>();
        }
        const cplx<T> base_n0 = viewval(1, 0);   // the mirror row's ky = 0 value
        __syncthreads();   // LDS is the transforms' scratch from here on
        unsigned omask = row == 0 ? (f == 0 ? 0x9u : 0x20u) : (3u << (2 * fk));
        if (ph.linear) omask = (row == 0 && f == 0) ? 0x3u : (omask & 0x3u);   // vx, vy only
        p1_outputs<T, NY>(a, row, base, base_n0, sn, kx_r, omask, a.store_phi_next, lds, tid);
        if (row == 0 && f == 1 && !ph.linear) {
            // outputs of row 0 that carry a factor kx = 0: vy, d/dx omega, d/dx rho
            const int p0 = g.nkxh - 1;
            for (int idx = tid; idx < 3 * NY; idx += TT) {
                const int z = idx / NY, y = idx - z * NY;
                const int o = z == 0 ? 1 : (z == 1 ? 2 : 4);
                a.G[(long long)o * g.I + ((long long)(y >> 2) * g.nkxh + p0) * 4 + (y & 3)] = czero<T>();
            }
        }
    }
}

}  // namespace shrt

// step_common.hpp -- data layout, control block and pointwise helpers shared by
// the step kernels and the auxiliary transforms.
//
// HBM layout (one instance):
//   state rings  a_omg[3], d_omg[3], a_rho[3], d_rho[3], aphi : nkx*nky complex,
//                reference layout f[ikx*nky + iky] (fields.cu:137-165)
//   G            6 fields x I complex, I = nkxh*ny; G[f][t][kx][c] holds the
//                y-inverse transform at (kx >= 0, y = 4t + c)
//   H            2 fields x I complex, same tiling, the x-forward transform of
//                the two brackets at (kx >= 0, y)
//   kx, ky       double tables (fourier.cu:123-128)
//   s[2][nkx]    FP64 shearing shift ky_shift (shear.cu:282), by step parity
//   jump[2][nkx] remap offsets of the last two remap events, by epoch parity
//   ctl[2]       StepCtl, by step parity
#pragma once

#include "fft.hpp"

namespace shrt {

constexpr int kTile = 4;   // rows (kx) per P1/P3 workgroup == columns (y) per P2 workgroup

// Waves per SIMD that the LDS of a kTile-team workgroup allows (160 KiB per
// CU, four SIMDs), capped at 4 (128 VGPRs).  Used as the occupancy target
// of the transform kernels so the compiler keeps their registers within it.
template <typename T, int N, int TT = TeamShape<T, N>::TT, int TEAMS = kTile>
constexpr int waves_per_eu() {
    constexpr long long lds = (long long)TEAMS * lds_len<N>() * (long long)sizeof(cplx<T>);
    constexpr long long wgs = (160LL * 1024LL) / lds;
    constexpr long long waves = wgs * ((TEAMS * TT + 63) / 64);
    constexpr long long w = waves / 4;
    return w < 1 ? 1 : (w > 4 ? 4 : (int)w);
}

// Points per thread of the P2 transforms: 16 from 2048 up in both precisions
// (FP64 2048^2: 77.5 us per launch against 87.2 us with the 8 points of its
// P1/P3 teams); the P1/P3 team shape below 2048 (4 points at FP64 1024, where
// the 256-thread columns alone took the step from 52.1 to 51.3-51.6 us; in FP32
// 4 points were slower, 35.8 against 35.1).
template <typename T, int N> constexpr int p2_points() { return N >= 2048 ? 16 : TeamShape<T, N>::P; }
// Columns (teams) per P2 workgroup: 1, 2 or kTile.  Two at FP64 2048 (P2
// 57.9 us per launch against 66.5 with one column); below 512, where a team
// is less than a wavefront, as many as fill one (at most kTile); one elsewhere
// (two columns measured 2048^2 FP32 28.7 against 29.0 us, FP32 1024 13.0
// against 12.1, FP64 1024 23.1 against 22.0, FP64 4096 341 against 311, a
// single 512^2 instance's step FP32 23.4 against 22.6 and FP64 33.9 against
// 32.0 though a batch of 89 such instances 5% faster; four columns at 2048^2
// FP32 31.5 us).
template <typename T, int N> constexpr int p2_teams() {
    constexpr int tt = Team<N, p2_points<T, N>()>::TT;
    if (sizeof(T) == 8 && N == 2048) return 2;
    if (tt < 64) return 64 / tt < kTile ? 64 / tt : kTile;
    return 1;
}

// Kernel variants chosen per context at run time (shrt_autotune): one code per
// pass.  Defaults: the fastest at 2048^2 FP32 (P1 31.3 us row against 32.4 us
// tile; P2 34.4 us three transforms against 36.8 us four; P3 29.7 us one row
// pair against 32.0 us four).
enum : int { kP1Row = 0, kP1Tile = 1, kP1Pair = 2 };   // per workgroup: one (row pair, output) / a tile of kTile
                                                       // row pairs / one (row pair, field group): two outputs
enum : int { kP2Three = 0, kP2Four = 1 };      // x-inverse transforms per column
enum : int { kP3Pair = 0, kP3Tile = 1, kP3Split = 2 };   // one row pair / kTile row pairs / one row of a pair
                                                         // per workgroup
struct Variants {
    int p1 = kP1Row, p2 = kP2Three, p3 = kP3Pair;
    int fuse = 0;   // captured batches run P3 of step n and P1 of step n+1 as one kernel
};

// Exchange synchronisation of a team: a wave-level fence for one-wave teams,
// otherwise a workgroup barrier.
template <int TT>
struct TeamSync {
    __device__ __forceinline__ void operator()() const {
        if constexpr (TT == 64) wave_sync();
        else __syncthreads();
    }
};

// Join the teams of a workgroup before a buffer shared between teams is
// rewritten.  Teams of several wavefronts already pass a workgroup barrier at
// the end of every transform; one-wave teams synchronise only their own wave.
template <int TT>
__device__ __forceinline__ void teams_join() {
    if constexpr (TT == 64) __syncthreads();
}

struct Geo {
    int nx, ny, nkx, nky, nkxh, nkxh2, nkxpad, ncy;
    int ntile_x;   // ceil(nkxh / kTile): row-pair tiles
    int ntile_y;   // ny / kTile: column tiles
    long long I;   // nkxh * ny
    double Lx, Ly, dx, dy;
    double scale;  // 1 / (nx*ny)  (fft.cu:165)
};

struct Phys {
    double nu, kappa, sigma, g, rho0, rho0_prime, cfl_num, eps;
    int linear;
    int narrow;   // FP32 contexts: 1 = the state update in FP32 arithmetic, rounded to nearest (the
                  // reference's FP32 build; its own P3 instance); 0 = evaluated in FP64 and rounded
                  // once, stochastically
};

// Device-resident control block.  Step n reads ctl_in and its last kernel
// writes ctl_out; the two alternate, so one captured graph serves every step.
struct StepCtl {
    double time;                       // t_n
    double delt;                       // delt before this step's CFL check
    long long istep;                   // n
    int epoch;                         // number of remaps before step n (frame of step n)
    int slot_epoch_a[3];               // frame of each a-ring slot (omega and rho)
    int slot_epoch_d[3];               // frame of each d-ring slot
    int phi_epoch;                     // frame of aphi
    int remap_count;
    int halvings;
    int error;                         // bit 0: remap while a live slot was still pending
    int log_count;
    long long last_remap_step;
};

struct HalvingRec {
    long long istep;
    double time, value, delt;
    int which, pad_;
    int counts[4];      // halvings per condition, in check_cfl's order
    double values[4];   // cfl_vx, cfl_vy, g/rho0, rho0_prime
    double delt_before;
};

template <typename T>
struct StepArgs {
    Geo g;
    Phys ph;
    cplx<T>* a_omg[3];
    cplx<T>* d_omg[3];
    cplx<T>* a_rho[3];
    cplx<T>* d_rho[3];
    cplx<T>* aphi;
    cplx<T>* G;
    cplx<T>* H;
    const double* kx;
    const double* ky;
    const cplx<T>* twx;
    const cplx<T>* twy;
    StepCtl* ctl_in;
    StepCtl* ctl_out;
    const double* s_in;
    double* s_out;
    int* jump;
    HalvingRec* hlog;
    int hlog_cap;
    int store_phi;        // this step's P1 stores aphi
    int store_phi_next;   // the fused next-step P1 (k_p3 with FUSE) stores aphi
    double* cflp;   // [ncflp][2]: per-workgroup max|vx + sigma*y|, max|vy| of P2
    int ncflp;      // P2 workgroups
};

// Wavefront max of a double.
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// Reduce P2's per-workgroup CFL maxima over the whole grid (every P3
// workgroup does this; `red` is workgroup-shared scratch of 2*16 doubles).
// One (max|vx + sigma*y|, max|vy|) pair per P2 workgroup in the working
// precision (the maxima of working-precision values, so exact), eight loads
// in flight per thread; NT = threads per workgroup.
template <typename T, int NT>
__device__ __forceinline__ void cfl_maxima(const double* __restrict__ cflp, int n, double* red, double* cvx,
                                           double* cvy) {
    const cplx<T>* p = reinterpret_cast<const cplx<T>*>(cflp);
    T m0 = 0, m1 = 0;
    int i = threadIdx.x;
    for (; i + 7 * NT < n; i += 8 * NT) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[i + u * NT];
#pragma unroll
        for (int u = 0; u < 8; ++u) { m0 = fmax(m0, v[u].x); m1 = fmax(m1, v[u].y); }
    }
    for (; i < n; i += NT) {
        const cplx<T> v = p[i];
        m0 = fmax(m0, v.x);
        m1 = fmax(m1, v.y);
    }
    m0 = wave_max(m0);
    m1 = wave_max(m1);
    constexpr int nw = (NT + 63) >> 6;
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w] = m0; red[16 + w] = m1; }
    __syncthreads();
    double r0 = 0.0, r1 = 0.0;
#pragma unroll
    for (int i = 0; i < nw; ++i) { r0 = fmax(r0, red[i]); r1 = fmax(r1, red[16 + i]); }
    *cvx = r0;
    *cvy = r1;
}

// Ring slots after the renew_fields rotation of step n (time_integral.cu:474-484):
// a0 (written) = (-(n+1)) mod 3, a1 (state n) = (-n) mod 3, a2 = (1-n) mod 3.
__host__ __device__ inline int slot_a0(long long n) { return (int)(((-(n + 1)) % 3 + 3) % 3); }
__host__ __device__ inline int slot_a1(long long n) { return (int)(((-n) % 3 + 3) % 3); }
__host__ __device__ inline int slot_a2(long long n) { return (int)(((1 - n) % 3 + 3) % 3); }

template <typename T> __device__ __forceinline__ cplx<T> czero() { return {T(0), T(0)}; }

// Work item of workgroup b in a grid of n.  The dispatcher deals consecutive
// workgroups to the eight XCDs in turn and each XCD has its own L2, so items
// are renumbered to give XCD x one contiguous range of items, in dispatch
// order: workgroups whose items share input rows then share an L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int b, int n) {
    const int x = b % kXcds, j = b / kXcds;
    const int base = n / kXcds, rem = n % kXcds;
    return x * base + min(x, rem) + j;
}

// A complex value over a real one: two IEEE divisions in FP64 (the reference's
// cmplx / real operator); in FP32 the hardware reciprocal refined by one Newton
// step and two products (within about 2 ulp; 0.6 µs less per 2048^2 step than
// an IEEE division).
template <typename T>
__device__ __forceinline__ cplx<T> cdivr(cplx<T> z, T d) {
    if constexpr (sizeof(T) == 8) {
        return {z.x / d, z.y / d};
    } else {
        float r = __builtin_amdgcn_rcpf(d);
        r = r * (2.0f - d * r);
        return {z.x * r, z.y * r};
    }
}

// Row stride of a workgroup staging buffer [team][y] between the team
// transforms (natural order) and the [t][team][c] tiles of G and H.  N + 8
// complex (== 8 mod 32) spreads the 16 (team, c) lanes of a tile row over
// distinct LDS banks; it fits the transform buffers from N = 256 up.
template <int N> constexpr int stage_stride() { return N >= 256 ? N + 8 : lds_len<N>(); }

// Row stride of P2's staging buffer [field][c][kx] for two G fields of a
// column tile: the kx extent rounded up to 32 plus 8 complex where that fits
// the transform buffers.
template <int NX> constexpr int p2_stage_stride() {
    constexpr int nk = NX / 3 + 1;
    constexpr int padded = (nk + 31) / 32 * 32 + 8;
    return 2 * padded <= lds_len<NX>() ? padded : nk;   // two fields per column
}

// Remapped read of a spectral buffer (shear.cu:296-322).  `jmp` is null when the
// buffer is already in the current frame.
template <typename T>
__device__ __forceinline__ cplx<T> view_load(const cplx<T>* __restrict__ buf, int row, int ky, const Geo& g,
                                             const int* __restrict__ jmp) {
    if (jmp == nullptr) return buf[(long long)row * g.nky + ky];
    const int j = jmp[row];
    if (j <= 0) {
        const int src = ky - j;
        return src < g.nky ? buf[(long long)row * g.nky + src] : czero<T>();
    }
    if (ky - j > 0) return buf[(long long)row * g.nky + (ky - j)];
    const int src = j - ky;
    const int mrow = row == 0 ? 0 : g.nkx - row;   // ikx_indexed (shear.cu:160-162)
    if (src >= g.nky) return czero<T>();
    const cplx<T> v = buf[(long long)mrow * g.nky + src];
    return {v.x, -v.y};
}

// Sheared ky and kperp2 in FP64 without contraction (shear.cu:335-336).
__device__ __forceinline__ double kys_of(const double* ky, const double* s, int row, int iky) {
    return add_rn(ky[iky], s[row]);
}
__device__ __forceinline__ double k2_of(double kx, double kys) {
    return add_rn(mul_rn(kx, kx), mul_rn(kys, kys));
}

// The step size after the CFL halvings of check_cfl (time_integral.cu:357-383).
// Returns the halving count; `which` gets the last condition that halved.
// `cnt` (may be null) receives the halvings per condition.
__device__ __forceinline__ int cfl_delt(const StepCtl* c, const Geo& g, const Phys& ph, double cvx, double cvy,
                                        double* dt_out, int* which, int* cnt = nullptr) {
    double dt = c->delt;
    const double cg = ph.g / ph.rho0;
    const double cr = ph.rho0_prime;
    int h = 0, h0 = 0, h1 = 0, h2 = 0;
    *which = -1;
    while (div_rn(mul_rn(cvx, dt), g.dx) > ph.cfl_num && h < 2200) { dt /= 2.0; ++h; *which = 0; }
    h0 = h;
    while (div_rn(mul_rn(cvy, dt), g.dy) > ph.cfl_num && h < 2200) { dt /= 2.0; ++h; *which = 1; }
    h1 = h;
    while (div_rn(mul_rn(cg, dt), g.dx) > ph.cfl_num && h < 2200) { dt /= 2.0; ++h; *which = 2; }
    h2 = h;
    while (div_rn(mul_rn(cr, dt), g.dx) > ph.cfl_num && h < 2200) { dt /= 2.0; ++h; *which = 3; }
    if (cnt) {
        cnt[0] = h0;
        cnt[1] = h1 - h0;
        cnt[2] = h2 - h1;
        cnt[3] = h - h2;
    }
    *dt_out = dt;
    return h;
}

__device__ __forceinline__ void sync_wg() { __syncthreads(); }

}  // namespace shrt

/* shrt.h -- C ABI of the MI355X-native shearing-rt pseudospectral step.
 *
 * The reference (Fulac/shearing-rt) has no plugin or FFI layer: its hot path is a
 * set of C++ functions on global state (time_integral.h, fft.h, shear.h,
 * fields.h) driven by main.cu.  Each entry point below replaces one of them; the
 * replaced reference interface is cited on every declaration.  Plain C types
 * only, no torch or HIP types.  The library owns device memory (HBM); the caller
 * owns host buffers, which use the reference's layouts:
 *
 *   spectral  f[ikx*nky + iky], complex interleaved {re, im}   (fields.cu:137-165)
 *             kx order 0..nkxh-1, -(nkxh-1)..-1                 (fourier.cu:124-125)
 *   physical  f[ix*ny + iy], real                              (fft.cu:160-164)
 *
 * The element type of host buffers is the context's precision ("cureal",
 * cmplx.h:7-13): float / float2 for SHRT_FP32, double / double2 for SHRT_FP64.
 * Precision is chosen per context at create time (it replaces -DDBLE).
 * Shift bookkeeping (ky_shift, remap trigger and offsets) and the time are FP64
 * in both precisions.
 *
 * Every call returns SHRT_OK (0) or an error code; nothing calls exit().  The
 * text of the last error of the calling thread is in shrt_last_error().  A
 * context is used from one host thread; its GPU work is ordered on its own
 * HIP stream.
 */
#ifndef SHRT_H
#define SHRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SHRT_OK = 0,
    SHRT_ERR_INVALID = 1,      /* bad argument or unsupported configuration        */
    SHRT_ERR_UNSUPPORTED = 2,  /* grid size outside the compiled transform lengths */
    SHRT_ERR_DEVICE = 3,       /* a HIP call failed                                */
    SHRT_ERR_ALLOC = 4,        /* device memory allocation failed                  */
    SHRT_ERR_STATE = 5         /* the device reported an error during a step       */
} shrt_status;

typedef enum { SHRT_FP32 = 4, SHRT_FP64 = 8 } shrt_precision;

/* Spectral fields (the reference's dv_aomg0 / dv_aphi / dv_arho0). */
typedef enum { SHRT_OMEGA = 0, SHRT_PHI = 1, SHRT_RHO = 2 } shrt_field;

/* Physical fields: omega, phi, rho (output_fields, file_access.cu:250-263), the
 * velocities vx, vy of the last CFL check (time_integral.cu:354), and the
 * advection term vx*d(omega)/dx + vy*d(omega)/dy of the current state that the
 * next step computes and output_pbk writes (poisson_bracket_shear,
 * shear.cu:395-434, called at time_integral.cu:500-505; the velocities are those
 * of phi = omega/kperp2 with the current shift, as check_cfl makes them). */
typedef enum {
    SHRT_P_OMEGA = 0, SHRT_P_PHI = 1, SHRT_P_RHO = 2, SHRT_P_VX = 3, SHRT_P_VY = 4, SHRT_P_BRACKET = 5
} shrt_pfield;

/* Ring buffers of the AB3/BDF2 history (fields.cu:21-25). */
typedef enum { SHRT_RING_AOMG = 0, SHRT_RING_DOMG = 1, SHRT_RING_ARHO = 2, SHRT_RING_DRHO = 3 } shrt_ring;

/* Parameters of input_data (file_access.cu:123-165), with the same meaning.
 * `delt` is the initial time step; check_cfl may halve it during a step. */
typedef struct {
    int nx, ny;
    double Lx, Ly;
    double delt;
    double cfl_num;
    double nu, kappa, sigma, g, rho0, rho0_prime, rho_eps1;
    int noise;        /* "initial noise" */
    int linear;       /* "linear eq"     */
    int precision;    /* SHRT_FP32 or SHRT_FP64 */
    int graph_steps;  /* steps per captured HIP graph (even, >= 2); 0 = launch eagerly */
} shrt_params;

/* Seeded initial condition (new; additive to the reference IC):
 * omega = 0; on retained modes with kmin <= |k| <= kmax,
 * rho_hat = amp * |k|^expo * exp(i*theta), theta from SplitMix64(seed). */
typedef struct {
    uint64_t seed;
    double kmin, kmax, amp, expo;
} shrt_seeded_ic;

typedef struct {
    int64_t istep;          /* number of completed steps (main.cu:14)          */
    double time;            /* simulated time (time_integral.cu:340)           */
    double delt;            /* current time step after CFL halvings            */
    int remap_count;        /* number of remap events (shear.cu:234)           */
    int halvings;           /* number of CFL halvings (time_integral.cu:361)   */
    int64_t last_remap_step;/* step index of the last remap, -1 if none        */
} shrt_time;

typedef struct {
    double energy;          /* 1/2 sum w kperp2_shear |phi_hat|^2  (FP64, compensated) */
    double enstrophy;       /* 1/2 sum w |omega_hat|^2                                   */
    double rho_var;         /* 1/2 sum w |rho_hat|^2                                     */
} shrt_diag;

typedef struct {
    int64_t istep;
    double time, value, delt;
    int which;              /* 0 cfl_vx, 1 cfl_vy, 2 g/rho0, 3 rho0_prime (the last that halved) */
    int count;              /* halvings taken in that step                */
    int counts[4];          /* halvings per condition, in check_cfl's order   */
    double values[4];       /* cfl_vx, cfl_vy, g/rho0, rho0_prime of the step */
    double delt_before;     /* delt entering the step; the printed lines of
                               time_integral.cu:361-383 follow by halving it  */
} shrt_halving;

typedef struct shrt_ctx shrt_ctx;

/* init_four + init_fields + init_tint + init_shear (main.cu:28-32; fourier.cu:116-150,
 * fields.cu:68-81, time_integral.cu:205-215, shear.cu:149-185).  All buffers zeroed. */
shrt_status shrt_create(const shrt_params* params, shrt_ctx** out);

/* finish_* (main.cu:69-73). */
shrt_status shrt_destroy(shrt_ctx* ctx);

/* initialize() (time_integral.cu:229-259): rho = rho_eps1*cos(2x+10y), omega = 0,
 * xtok, optional glibc-rand() noise (init_dis, :261-320), aphi = omega/kperp2. */
shrt_status shrt_initialize(shrt_ctx* ctx);

/* Seeded initial condition (new). */
shrt_status shrt_initialize_seeded(shrt_ctx* ctx, const shrt_seeded_ic* ic);

/* time_advance(istep, time) (time_integral.cu:324-341), `nsteps` times.  Returns
 * after the device has finished and reports device-side errors. */
shrt_status shrt_step(shrt_ctx* ctx, int nsteps);

/* Enqueue `nsteps` steps on the context's stream and return at once (no sync,
 * no error check); pair with shrt_synchronize(). */
shrt_status shrt_step_async(shrt_ctx* ctx, int nsteps);
shrt_status shrt_synchronize(shrt_ctx* ctx);

/* Read / write the current spectral field (dv_aomg0, dv_aphi, dv_arho0), host
 * buffer of nkx*nky complex.  Setting omega also sets aphi = omega/kperp2_shear
 * (as initialize() does, time_integral.cu:257). */
shrt_status shrt_get_spectral(shrt_ctx* ctx, int field, void* host);
shrt_status shrt_set_spectral(shrt_ctx* ctx, int field, const void* host);

/* The current spectral field restricted to the retained modes of a coarser
 * nx_c x ny_c grid (new; the spectral filter of the learned closure's training
 * data): host buffer of nkx_c*nky_c complex in the coarse grid's layout, row i
 * from this grid's row i (kx >= 0) or nkx - (nkx_c - i) (kx < 0), its first
 * nky_c modes; the same physical wavenumbers when Lx and Ly are equal.  Only
 * those modes cross to the host. */
shrt_status shrt_get_spectral_modes(shrt_ctx* ctx, int field, int nx_c, int ny_c, void* host);

/* The spectral fields as update_shear hands them to k_data_bef at a remap step
 * (shear.cu:236, before the shearing_field calls of :238-268): right after a
 * step that remapped, omega and rho of the new state and the aphi of that
 * step, all in the frame the step started in; at any other time the same as
 * shrt_get_spectral.  The remap itself is applied lazily: the stored buffers
 * keep the old frame until the next step reads them through the remap. */
shrt_status shrt_get_spectral_unremapped(shrt_ctx* ctx, int field, void* host);

/* Ring access by the reference's names after the last rotation: slot 0 = *0,
 * 1 = *1, 2 = *2 (renew_fields, time_integral.cu:474-484). */
shrt_status shrt_get_buffer(shrt_ctx* ctx, int ring, int slot, void* host);
shrt_status shrt_set_buffer(shrt_ctx* ctx, int ring, int slot, const void* host);

/* Physical field of omega, phi, rho or the velocities, host buffer of nx*ny
 * reals.  lab_frame = 0: ktox (fft.cu:170-204) in the sheared frame; lab_frame
 * = 1: what output_fields writes (file_access.cu:250-259), ktox_shear
 * (shear.cu:467-478) with the current ky_shift when sigma != 0, ktox otherwise.
 * The velocities (those of the last CFL check, time_integral.cu:354) and the
 * bracket exist in the sheared frame only. */
shrt_status shrt_get_physical(shrt_ctx* ctx, int pfield, int lab_frame, void* host);

/* xtok / ktox (fft.h:17-25) on host buffers. */
shrt_status shrt_xtok(shrt_ctx* ctx, const void* phys_in, void* spec_out);
shrt_status shrt_ktox(shrt_ctx* ctx, const void* spec_in, void* phys_out);

/* Allocation-free xtok on device buffers.  No counterpart in the reference: for
 * callers that cannot allocate device memory at the call (HIP graph capture,
 * serving loops).  phys_dev (nx*ny reals) and spec_dev (nkx*nky complex) are the
 * caller's device buffers; the result equals shrt_xtok's.  The x-forward
 * intermediate (nkxh*ny complex) lives in a 16 MiB scratch array of the library
 * (enough for 2048x2048 in FP32, 2048x1024 or 1024x2048 in FP64; a grid whose
 * intermediate is larger is refused with SHRT_ERR_INVALID) that every GPU the library runs on holds until the process
 * ends.  Enqueued on the context's stream without synchronising (pair with
 * shrt_synchronize); no memset, no allocation, no host wait.  Calls that overlap
 * on one GPU, from any contexts, share the scratch and race. */
shrt_status shrt_xtok_noalloc(shrt_ctx* ctx, const void* phys_dev, void* spec_dev);

/* ktox_shear (shear.cu:467-478) on a host buffer with the context's current
 * ky_shift: the lab-frame physical field of a spectral field. */
shrt_status shrt_ktox_shear(shrt_ctx* ctx, const void* spec_in, void* phys_out);

/* shearing_field (shear.cu:296-322): remap one spectral buffer with the given
 * jump[nkx] (ikx_indexed fold for jump > 0). */
shrt_status shrt_shearing_field(shrt_ctx* ctx, const int* jump, const void* in, void* out);

/* Time and shift bookkeeping: time / istep / delt, and ky_shift[nkx] (FP64) plus
 * jump[nkx] (shear.cu:21-23) as shearing_ky leaves it after the last step
 * (shear.cu:275-294): round(ky_shift/kx[1]) before a remap zeroed the shift,
 * i.e. the remap's table right after a remap step and the current shift's
 * otherwise; all zero before the first step. */
shrt_status shrt_get_time(shrt_ctx* ctx, shrt_time* out);
shrt_status shrt_set_time(shrt_ctx* ctx, int64_t istep, double time, double delt);
shrt_status shrt_get_shift(shrt_ctx* ctx, double* ky_shift, int* jump);
shrt_status shrt_set_shift(shrt_ctx* ctx, const double* ky_shift);

/* FP64 compensated energy / enstrophy / density variance (new). */
shrt_status shrt_diagnostics(shrt_ctx* ctx, shrt_diag* out);

/* CFL halving events recorded on the device (the printf lines of check_cfl,
 * time_integral.cu:363-382).  Returns the number copied (<= cap) in *count. */
shrt_status shrt_get_halvings(shrt_ctx* ctx, shrt_halving* out, int cap, int* count);

/* Benchmark support (no reference counterpart).  shrt_time_steps enqueues
 * `nsteps` steps exactly as shrt_step_async does (graph replays when
 * graph_steps > 0) between two HIP events on the context's stream and returns
 * the elapsed milliseconds.  shrt_kernel_times launches `nsteps` steps kernel
 * by kernel, each dispatch recording HIP events at its own start and end
 * (hipExtLaunchKernelGGL, so launch gaps are excluded), and returns the
 * average duration in milliseconds of each of the three step kernels
 * (ms[0..2] = y-inverse/Poisson pass, x pass with brackets, y-forward/update
 * pass). */
shrt_status shrt_time_steps(shrt_ctx* ctx, int nsteps, double* ms_total);
shrt_status shrt_kernel_times(shrt_ctx* ctx, int nsteps, double* ms3);
/* The fused kernel's duration (no reference counterpart): `nsteps` + 1 steps
 * launched as a fused captured batch launches them (P1; P2 and the fused P3 of
 * a step with P1 of the next, per step; P2 and P3 last), each fused dispatch
 * recording events at its own start and end; *ms = its average duration in
 * milliseconds.  Refused for the FP32-arithmetic update, which runs unfused. */
shrt_status shrt_fused_kernel_time(shrt_ctx* ctx, int nsteps, double* ms);

/* Benchmark support (no reference counterpart): copy `bytes` (a multiple of 16)
 * between two fresh device buffers with a plain copy kernel, `reps` times, and
 * return the average milliseconds per copy.  Its HBM counters are the
 * reference point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings. */
shrt_status shrt_copy_bench(long long bytes, int reps, double* ms);

/* Benchmark support (no reference counterpart): the same measurement with the
 * fastest copy kernel found on MI355X (four 16-byte words in flight per
 * thread, nontemporal loads and stores), i.e. the achievable HBM copy rate. */
shrt_status shrt_stream_copy_bench(long long bytes, int reps, double* ms);

/* Autotuning (new; no reference counterpart).  Each pass of the step has
 * kernel variants: P1 one (row pair, output) per workgroup (0), a tile of four
 * row pairs (1) or one (row pair, field group) with its two outputs (2); P2
 * three x-inverse transforms per column (0) or four (1); P3 one row pair per
 * workgroup (0), four (1) or one row of a pair (2: twice the workgroups, each
 * transforming the pair's H row and updating its own row).  shrt_autotune times the variants of each pass on
 * a scratch instance of the context's shape and parameters (reps launches
 * each) and keeps the fastest.  The variants agree to rounding (the P2
 * variants pair the fields differently in the complex transforms; all are
 * compiled separately).  ms[] = average ms per launch of p1 row, p1 tile,
 * p1 pair, p2 three, p2 four, p3 pair, p3 tile, p3 split.  It then times
 * captured batches of graph_steps steps (20 if the context has none) with and
 * without the P3+P1 fusion (shrt_set_fusion) and keeps the faster.  Without
 * autotuning a context starts with P1 variant 2 for ny = 2048 in FP32 and for
 * ny >= 4096, 1 for ny = 512, 0 otherwise (P2 0, P3 0), and with the P3+P1
 * fusion on for FP32 at ny = 512, 1024 and 2048, off otherwise: the picks of
 * shrt_autotune on MI355X.  shrt_get_variants reports the current choice. */
typedef struct {
    int p1, p2, p3;
    double ms[8];
    int fuse;            /* 1: captured batches fuse P3 with the next P1 (shrt_set_fusion) */
    double step_ms[2];   /* ms per step of captured batches, unfused and fused */
} shrt_tuning;
shrt_status shrt_autotune(shrt_ctx* ctx, int reps, shrt_tuning* out);
shrt_status shrt_set_variants(shrt_ctx* ctx, int p1, int p2, int p3);
shrt_status shrt_get_variants(shrt_ctx* ctx, int* p1, int* p2, int* p3, int* fuse);

/* Precision of the FP32 state update (new; the reference's choice is its build
 * flag -DDBLE).  wide = 1 (the default): the AB/BDF2 combination and the
 * implicit-diffusion division (time_integral.cu:633-695, 747-767) are evaluated
 * in FP64 from the FP32 operands and rounded once, stochastically (unbiased,
 * deterministic: keyed by step, field and mode).  wide = 0: the same formulas in
 * FP32 arithmetic rounded to nearest, with the coefficients the reference's FP32
 * build forms in double ((2.0/3.0)*delt, delt/9.0, 1 + c*delt*nu*k^2), i.e. that
 * build's arithmetic; it runs as its own P3 kernel, one row pair per workgroup,
 * without the P3+P1 fusion (wide = 0 turns fusion off).  FP64 contexts ignore
 * the flag.  Set it before an ensemble that holds the context is created. */
shrt_status shrt_set_update_precision(shrt_ctx* ctx, int wide);

/* Step fusion (new; no reference counterpart).  With on = 1 a captured batch
 * of graph_steps steps runs P3 of step n and P1 of step n+1 as one kernel: the
 * updated rows stay in LDS, the next shift and remap are formed in the kernel,
 * and the next step's y-inverse transforms are written without reading the new
 * state back.  on = 0 runs the three passes per step; a context starts with
 * fusion on for FP32 at ny = 512, 1024 and 2048, off otherwise.  Single steps
 * (shrt_step with fewer than graph_steps left) never fuse.  Fusion saves P1's
 * reads and one launch per step but pays on some shapes only; shrt_autotune
 * times both and keeps the faster. */
shrt_status shrt_set_fusion(shrt_ctx* ctx, int on);

/* Ensembles (new; no reference counterpart: the reference runs one instance
 * per process).  Independent contexts -- any sizes, precisions, parameters and
 * seeds -- are advanced together: each full batch of steps_per_graph steps (even)
 * is one captured HIP graph holding every member's step chain, so small
 * instances fill the GPU together.  Members keep their own state and API; a
 * member must not be destroyed before its ensemble.  The results equal the
 * members' individual runs bit for bit. */
typedef struct shrt_ensemble shrt_ensemble;
shrt_status shrt_ensemble_create(shrt_ctx* const* members, int n, int steps_per_graph, shrt_ensemble** out);
shrt_status shrt_ensemble_destroy(shrt_ensemble* ens);
/* nsteps steps of every member (full graphs, then single steps per member). */
shrt_status shrt_ensemble_step(shrt_ensemble* ens, int nsteps);
shrt_status shrt_ensemble_synchronize(shrt_ensemble* ens);
/* Timed run of nsteps (a multiple of steps_per_graph) between HIP events. */
shrt_status shrt_ensemble_time_steps(shrt_ensemble* ens, int nsteps, double* ms_total);

/* Learned closure (new; no reference counterpart).  A spectral-convolution
 * network on an n x n grid (n a power of two, 8..512): inputs [batch][4][n*n]
 * (channels omega, rho, phi, shear rate; physical layout f[ix*n + iy]),
 * outputs [batch][2][n*n] (corrections of the omega and rho tendencies):
 *   h0 = P x + p;  h_{l+1} = gelu(W_l h_l + w_l + K_l h_l), l < layers;
 *   out = Q h_L + q,
 * where K_l multiplies each retained mode (|kx| <= kmax, 0 <= ky <= kmax, half
 * plane) of its input by a complex width x width matrix.  The truncated
 * transforms, the mode mixing and the pointwise layers are GEMMs on the
 * matrix cores with BF16 operands and FP32 accumulation; training (MSE loss,
 * Adam) runs on the device on a data set uploaded once.  Parameters are one
 * FP32 array: P[width][4], p[width]; per layer a[modes][width][width] (real
 * parts, [out][in]), b[...] (imaginary parts), W[width][width], w[width]; then
 * Q[2][width], q[2].  Modes are ordered ky = 0 with kx = 0..kmax, then each
 * ky = 1..kmax with kx = -kmax..kmax. */
typedef struct shrt_closure shrt_closure;
shrt_status shrt_closure_create(int n, int kmax, int width, int layers, int max_batch, uint64_t seed,
                                shrt_closure** out);
shrt_status shrt_closure_destroy(shrt_closure* net);
shrt_status shrt_closure_info(shrt_closure* net, long long* nparams, int* nmodes);
shrt_status shrt_closure_get_params(shrt_closure* net, float* params);
/* Sets the parameters and resets the optimiser state. */
shrt_status shrt_closure_set_params(shrt_closure* net, const float* params);
shrt_status shrt_closure_forward(shrt_closure* net, const float* in, int batch, float* out);
/* Loss mean((out - target)^2) over batch*2*n*n values and its gradient with
 * respect to every parameter (no update). */
shrt_status shrt_closure_gradients(shrt_closure* net, const float* in, const float* target, int batch, double* loss,
                                   float* grad);
/* Upload a training set: in [nsamples][4][n*n], target [nsamples][2][n*n]. */
shrt_status shrt_closure_set_data(shrt_closure* net, const float* in, const float* target, int nsamples);
/* iters Adam steps (beta 0.9/0.999) on batches drawn with replacement from the
 * training set by a seeded generator; losses[iters] receives each batch loss
 * before its update. */
shrt_status shrt_closure_train(shrt_closure* net, int iters, int batch, float lr, uint64_t seed, double* losses);

/* In-place correction of a solver context by the network, all on the device
 * (new; no reference counterpart).  Inputs: ktox of omega, rho and the stored
 * phi of ctx (sheared frame), each divided by scale_in[i], and sigma as the
 * fourth channel; the outputs times scale_out[j] are transformed with xtok,
 * restricted to |kx|, ky <= kcut when kcut > 0 (the network's pointwise path
 * makes every scale; the correction is kept to the resolved band it is
 * trained for), and added, times dt_tot, to omega and rho (phi follows omega
 * as in shrt_set_spectral).  Fallback: nothing changes, and *applied = 0, when an
 * output or a correction is non-finite or |correction| * dt_tot > guard *
 * |field| (L2 over the stored modes) for either field.  The network's grid must be ctx's (square).
 * Same arithmetic as doing these steps through the host API. */
shrt_status shrt_closure_correct(shrt_closure* net, shrt_ctx* ctx, const double* scale_in, double sigma,
                                 const double* scale_out, double dt_tot, double guard, int kcut, int* applied);

/* Sizes derived by init_fft (fft.cu:91-96): out[0..7] = nkx, nky, nkxh, nkxh2,
 * nkxpad, ncy, nkx*nky, nkxh*ny. */
shrt_status shrt_sizes(int nx, int ny, int* out8);

/* Whether the library can run a grid: 1 if nx, ny are supported powers of two. */
int shrt_supported(int nx, int ny, int precision);

/* Number of visible HIP devices (0 when none or the runtime is unavailable).
 * Library utility with no reference counterpart. */
int shrt_device_count(void);

/* Last error message of the calling thread ("" if none). */
const char* shrt_last_error(void);

/* Library version string. */
const char* shrt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SHRT_H */

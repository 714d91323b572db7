// shrt_main.cpp -- the shrt driver: the reference's main.cu loop (main.cu:13-76)
// on the MI355X library, reading ./config.ini like input_data
// (file_access.cu:123-165) and writing the reference's output files.
//
//   shrt [--config PATH] [--dble] [--graph-steps N] [--print-config]
//
// --dble selects FP64 (the reference's `make dble=on`); the default is FP32 as
// in the reference's default build.  The loop variables (time, the next
// output times, tmax) have the build's precision as cureal does, so an FP32
// run takes the reference's float-accumulated step count and file names.
// Between events the driver runs batches of steps on the device (one captured
// HIP graph per graph-steps steps); a batch ends on the step at which the
// loop body could next act (an output time, the loop test, a remap), found by
// replaying the loop's arithmetic with the current delt: delt only ever
// halves, which moves every event later.  The CFL lines the reference prints
// inside check_cfl (time_integral.cu:361-383) are printed after each batch
// from the device's halving log, in order, which also gives each step's delt
// for the time accumulation.  An energy/enstrophy series is appended to
// energy.dat at every field output (new; additive).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shrt.h"
#include "ini.hpp"
#include "writers.hpp"

using namespace shrt_cli;

namespace {

struct Config {
    int nx, ny, nthread;
    double Lx, Ly, delt, tmax, cfl_num;
    double output_time;
    int nwrite;
    bool write_fields;
    bool noise_flag, linear_flag;
    double nu, kappa, sigma, g, rho0_prime, rho0, rho_eps1, rho_eps2;
};

// input_data (file_access.cu:123-165): the same keys, defaults and order of
// warnings.  With FP32 the floating values are read as float, as cureal is.
Config read_config(const Ini& pt, bool dble) {
    auto real = [&](const char* sec, const char* key, double def) -> double {
        return dble ? pt.entry_double(sec, key, def, nullptr) : (double)pt.entry_float(sec, key, (float)def, nullptr);
    };
    Config c;
    c.nx = pt.entry_int("simulation", "nx", 512, nullptr);
    c.ny = pt.entry_int("simulation", "ny", 1024, nullptr);
    c.nthread = pt.entry_int("simulation", "cuda thread num", 1024, nullptr);
    c.Lx = real("simulation", "Lx", M_PI);
    c.Ly = real("simulation", "Ly", M_PI);
    c.delt = real("simulation", "time step", 1e-3);
    c.tmax = real("simulation", "time max", 30);
    c.cfl_num = real("simulation", "cfl number", 1e-1);
    c.output_time = real("output", "output time step", 1.0);
    c.nwrite = pt.entry_int("output", "output loop count", 100, nullptr);
    c.write_fields = pt.entry_bool("output", "write output", true, nullptr);
    c.noise_flag = pt.entry_bool("problem", "initial noise", true, nullptr);
    c.linear_flag = pt.entry_bool("problem", "linear eq", false, nullptr);
    c.nu = real("problem", "nu", 1e-3);
    c.kappa = real("problem", "kappa", 1e-5);
    c.sigma = real("problem", "sigma", 1.0);
    c.g = real("problem", "g", 1.0);
    c.rho0_prime = real("problem", "rho0_prime", 1.0);
    c.rho0 = real("problem", "rho0", 1.0);
    c.rho_eps1 = real("problem", "rho_eps1", 1e-2);
    c.rho_eps2 = 0.1 * c.rho_eps1;
    return c;
}

void print_config(const Config& c, bool dble) {
    std::printf("nx = %d\nny = %d\ncuda thread num = %d\nLx = %.17g\nLy = %.17g\ntime step = %.17g\n"
                "time max = %.17g\ncfl number = %.17g\noutput time step = %.17g\noutput loop count = %d\n"
                "write output = %d\ninitial noise = %d\nlinear eq = %d\nnu = %.17g\nkappa = %.17g\n"
                "sigma = %.17g\ng = %.17g\nrho0_prime = %.17g\nrho0 = %.17g\nrho_eps1 = %.17g\n"
                "rho_eps2 = %.17g\nprecision = %s\n",
                c.nx, c.ny, c.nthread, c.Lx, c.Ly, c.delt, c.tmax, c.cfl_num, c.output_time, c.nwrite,
                (int)c.write_fields, (int)c.noise_flag, (int)c.linear_flag, c.nu, c.kappa, c.sigma, c.g,
                c.rho0_prime, c.rho0, c.rho_eps1, c.rho_eps2, dble ? "fp64" : "fp32");
}

void check(shrt_status s, const char* what) {
    if (s != SHRT_OK) {
        std::fprintf(stderr, "shrt: %s failed: %s\n", what, shrt_last_error());
        std::exit(1);
    }
}

Grid make_grid(const Config& c) {
    int sz[8];
    check(shrt_sizes(c.nx, c.ny, sz), "shrt_sizes");
    Grid g;
    g.nx = c.nx;
    g.ny = c.ny;
    g.nkx = sz[0];
    g.nky = sz[1];
    g.nkxh = sz[2];
    const double dx = c.Lx / c.nx, dy = c.Ly / c.ny;
    g.xx.resize(c.nx + 1);
    g.yy.resize(c.ny + 1);
    for (int ix = 0; ix <= c.nx; ++ix) g.xx[ix] = ix * dx;   // time_integral.cu:236-237
    for (int iy = 0; iy <= c.ny; ++iy) g.yy[iy] = iy * dy;
    g.kx.assign(g.nkx, 0.0);
    g.ky.assign(g.nky, 0.0);
    for (int ikx = 0; ikx < g.nkxh; ++ikx) g.kx[ikx] = ikx * 2 * M_PI / c.Lx;   // fourier.cu:123-128
    for (int ikx = 0; ikx < g.nkxh - 1; ++ikx) g.kx[g.nkx - ikx - 1] = -g.kx[ikx + 1];
    for (int iky = 0; iky < g.nky; ++iky) g.ky[iky] = iky * 2 * M_PI / c.Ly;
    g.kx_index.resize(g.nkx);   // file_access.cu:210-212
    for (int ikx = g.nkxh; ikx < g.nkx; ++ikx) g.kx_index[ikx - g.nkxh] = ikx;
    for (int ikx = 0; ikx < g.nkxh; ++ikx) g.kx_index[ikx + (g.nkx - g.nkxh)] = ikx;
    return g;
}

template <typename R>
struct Driver {
    shrt_ctx* ctx;
    const Config& c;
    const Grid& g;
    std::vector<R> ao, ap, ar;             // spectral, interleaved complex
    std::vector<R> omg, phi, rho, vx, vy;  // physical
    std::vector<R> pbk;                    // advection term of the step about to run
    int printed_halvings = 0;
    bool energy_started = false;
    // The loop variables of main.cu:14-20 in the build's precision (cureal):
    // float in FP32 runs, where the reference accumulates time in float and so
    // decides the step count, the output steps and the file names with it.
    int istep = 0;
    R time = 0, tmax = 0, next_output_time = 0, next_output_maxamp_time = 0;
    double delt = 0;                       // exact in R: read as R, only ever halved

    Driver(shrt_ctx* x, const Config& cc, const Grid& gg) : ctx(x), c(cc), g(gg) {
        const size_t M = (size_t)g.nkx * g.nky, N = (size_t)g.nx * g.ny;
        ao.resize(2 * M); ap.resize(2 * M); ar.resize(2 * M);
        omg.resize(N); phi.resize(N); rho.resize(N); vx.resize(N); vy.resize(N);
        pbk.resize(N);
    }

    shrt_time now() {
        shrt_time t;
        check(shrt_get_time(ctx, &t), "shrt_get_time");
        return t;
    }

    void spectra() {
        check(shrt_get_spectral(ctx, SHRT_OMEGA, ao.data()), "shrt_get_spectral");
        check(shrt_get_spectral(ctx, SHRT_PHI, ap.data()), "shrt_get_spectral");
        check(shrt_get_spectral(ctx, SHRT_RHO, ar.data()), "shrt_get_spectral");
    }

    // ktox or ktox_shear of omega, phi, rho (output_fields, output_maxamp)
    void physical() {
        check(shrt_get_physical(ctx, SHRT_P_OMEGA, 1, omg.data()), "shrt_get_physical");
        check(shrt_get_physical(ctx, SHRT_P_PHI, 1, phi.data()), "shrt_get_physical");
        check(shrt_get_physical(ctx, SHRT_P_RHO, 1, rho.data()), "shrt_get_physical");
    }

    void output_fields_now(int istep, double time) {
        physical();
        check(shrt_get_physical(ctx, SHRT_P_VX, 0, vx.data()), "shrt_get_physical");
        check(shrt_get_physical(ctx, SHRT_P_VY, 0, vy.data()), "shrt_get_physical");
        output_fields<R>(g, istep, c.nwrite, time, omg.data(), phi.data(), rho.data(), vx.data(), vy.data());
        shrt_diag d;
        check(shrt_diagnostics(ctx, &d), "shrt_diagnostics");
        energy_series(!energy_started, time, d.energy, d.enstrophy, d.rho_var);
        energy_started = true;
    }

    void output_maxamp_now(double time) {
        physical();
        const size_t N = omg.size();
        output_maxamp<R>(time, max_abs(omg.data(), N), max_abs(phi.data(), N), max_abs(rho.data(), N));
    }

    void spectra_unremapped() {
        check(shrt_get_spectral_unremapped(ctx, SHRT_OMEGA, ao.data()), "shrt_get_spectral_unremapped");
        check(shrt_get_spectral_unremapped(ctx, SHRT_PHI, ap.data()), "shrt_get_spectral_unremapped");
        check(shrt_get_spectral_unremapped(ctx, SHRT_RHO, ar.data()), "shrt_get_spectral_unremapped");
    }

    void k_data_bef_now(double time, int istep) {
        spectra();
        k_data_bef<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
    }

    // Steps to run before the loop body of main.cu can see an event, found by
    // replaying the loop's own arithmetic (time += delt in R, the FP64 shift
    // recurrence and remap test of shearing_ky) with the current delt.  A CFL
    // halving inside the batch only shortens delt, which makes every event
    // come later, never earlier.  Events: the loop test (time <= tmax), field
    // and max-amplitude output times, the step before one whose bracket is
    // written (advect_omg's time + delt > next_output_time), the k_data_bef of
    // istep 1, and a remap (k_data_bef / k_data_aft inside update_shear).
    long long plan(const std::vector<double>& shift) const {
        const long long kcap = 1000000;
        if (istep == 0) return 1;
        const bool remaps = c.sigma != 0 && g.nkx > 1;
        const double sdt1 = remaps ? (c.sigma * g.kx[1]) * delt : 0.0;
        double s1 = remaps ? shift[1] : 0.0;
        R t = time;
        const R dt = (R)delt;
        for (long long j = 0; j < kcap; ++j) {
            t = t + dt;
            if (remaps) {
                s1 = s1 - sdt1;
                if (std::fabs(s1 - (-g.kx[1])) < 1e-10) return j + 1;
            }
            if (!(t <= tmax)) return j + 1;
            if (c.write_fields && (t > next_output_time || t > next_output_maxamp_time)) return j + 1;
            if (c.write_fields && !c.linear_flag && t + dt > next_output_time) return j + 1;
        }
        return kcap;
    }

    // The step sizes of the steps [from, to) from the device's halving log:
    // delt carries over from step to step and halves at the logged steps.
    // Prints the CFL lines of check_cfl (time_integral.cu:361-383) with the
    // time the loop holds at each of them, and returns the loop's time after
    // the steps.
    R advance_time(long long from, long long to, double& dt) {
        std::vector<shrt_halving> h(4096);
        int n = 0;
        check(shrt_get_halvings(ctx, h.data(), (int)h.size(), &n), "shrt_get_halvings");
        static const char* names[4] = {"cfl_vx", "cfl_vy", "cfl_(g/rho0)", "cfl_rho0_prime"};
        int next = printed_halvings;
        R t = time;
        for (long long i = from; i < to; ++i) {
            if (next < n && h[next].istep == i) {
                double d = h[next].delt_before;
                for (int k = 0; k < 4; ++k)
                    for (int j = 0; j < h[next].counts[k]; ++j) {
                        d /= 2.0;
                        std::printf("istep = %d, time = %g, %s = %g, delt = %g\n", (int)i, (double)t, names[k],
                                    h[next].values[k], d);
                    }
                dt = d;
                ++next;
            }
            t = t + (R)dt;
        }
        printed_halvings = next;
        return t;
    }

    int run() {
        shrt_time st = now();
        istep = (int)st.istep;
        delt = st.delt;
        time = (R)st.time;
        tmax = (R)c.tmax;
        next_output_time = (R)c.output_time;          // file_access.cu:195
        const R output_maxamp_time = (R)0.1;          // main.cu:18-19
        next_output_maxamp_time = output_maxamp_time;
        int remaps_seen = st.remap_count;
        std::vector<double> shift(g.nkx);

        std::printf("yy[ny] = %lf\n", g.yy[g.ny]);
        if (c.write_fields) {
            output_fields_now(istep, time);
            output_maxamp_now(time);
            spectra();
            en_spectral<R>(g, istep, time, ao.data(), ap.data(), ar.data());
            ks_reim<R>(g, time, ao.data(), ap.data(), ar.data());
            k_data_bef<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
            k_data_aft<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
        }

        while (time <= tmax) {
            // advect_omg (time_integral.cu:500-505) writes the bracket it computes
            // when the step ends past the next output time; check_cfl can only
            // shorten delt, so a step that may end there runs alone and takes the
            // bracket of the current state first.
            const bool pbk_now = c.write_fields && !c.linear_flag;
            const bool pbk_due = pbk_now && time != 0 && time + (R)delt > next_output_time;
            if (pbk_now && time == 0) output_pbk<R>(g, 0.0, pbk.data());
            long long k = 1;
            if (pbk_due) {
                check(shrt_get_physical(ctx, SHRT_P_BRACKET, 0, pbk.data()), "shrt_get_physical");
            } else {
                if (c.sigma != 0) check(shrt_get_shift(ctx, shift.data(), nullptr), "shrt_get_shift");
                k = plan(shift);
            }
            const long long from = istep;
            check(shrt_step(ctx, (int)k), "shrt_step");
            st = now();
            istep = (int)st.istep;
            time = advance_time(from, st.istep, delt);
            if (pbk_due && time > next_output_time) output_pbk<R>(g, time, pbk.data());

            // update_shear's dumps around a remap (shear.cu:236, 270), with the
            // time and step after the step; the batch ends on the remap step
            if (st.remap_count != remaps_seen) {
                if (st.remap_count == remaps_seen + 1 && st.last_remap_step == st.istep - 1 && c.write_fields) {
                    spectra_unremapped();
                    k_data_bef<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
                    spectra();
                    k_data_aft<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
                } else if (c.write_fields) {
                    std::fprintf(stderr, "shrt: remap at step %lld was not the last of its batch; befk/aftk not written\n",
                                 (long long)st.last_remap_step);
                }
                remaps_seen = st.remap_count;
            }

            if (istep == 1) k_data_bef_now(time, istep);
            if (c.write_fields && time > next_output_time) {
                next_output_time += (R)c.output_time;
                output_fields_now(istep, time);
                spectra();
                en_spectral<R>(g, istep, time, ao.data(), ap.data(), ar.data());
                ks_reim<R>(g, time, ao.data(), ap.data(), ar.data());
                if (c.linear_flag || c.sigma == 0) k_data_bef<R>(g, time, istep, c.nwrite, ao.data(), ap.data(), ar.data());
            }
            if (c.write_fields && time > next_output_maxamp_time) {
                next_output_maxamp_time += output_maxamp_time;
                output_maxamp_now(time);
            }
        }
        return 0;
    }
};

}  // namespace

int main(int argc, char** argv) {
    std::string path = "config.ini";
    bool dble = false, print_only = false;
    int graph_steps = 20;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--config") && i + 1 < argc) path = argv[++i];
        else if (!std::strcmp(argv[i], "--dble")) dble = true;
        else if (!std::strcmp(argv[i], "--print-config")) print_only = true;
        else if (!std::strcmp(argv[i], "--graph-steps") && i + 1 < argc) graph_steps = std::atoi(argv[++i]);
        else {
            std::fprintf(stderr, "usage: shrt [--config PATH] [--dble] [--graph-steps N] [--print-config]\n");
            return 2;
        }
    }
    Ini pt;
    try {
        pt = Ini::read(path);
    } catch (const std::exception& e) {
        std::printf("ERROR: unable to read config file: %s", e.what());
        return 1;
    }
    const Config c = read_config(pt, dble);
    if (print_only) {
        print_config(c, dble);
        return 0;
    }

    shrt_params p{};
    p.nx = c.nx;
    p.ny = c.ny;
    p.Lx = c.Lx;
    p.Ly = c.Ly;
    p.delt = c.delt;
    p.cfl_num = c.cfl_num;
    p.nu = c.nu;
    p.kappa = c.kappa;
    p.sigma = c.sigma;
    p.g = c.g;
    p.rho0 = c.rho0;
    p.rho0_prime = c.rho0_prime;
    p.rho_eps1 = c.rho_eps1;
    p.noise = c.noise_flag ? 1 : 0;
    p.linear = c.linear_flag ? 1 : 0;
    p.precision = dble ? SHRT_FP64 : SHRT_FP32;
    p.graph_steps = graph_steps;
    shrt_ctx* ctx = nullptr;
    check(shrt_create(&p, &ctx), "shrt_create");
    check(shrt_initialize(ctx), "shrt_initialize");
    const Grid g = make_grid(c);
    int rc;
    if (dble) {
        Driver<double> d(ctx, c, g);
        rc = d.run();
    } else {
        Driver<float> d(ctx, c, g);
        rc = d.run();
    }
    check(shrt_destroy(ctx), "shrt_destroy");
    return rc;
}

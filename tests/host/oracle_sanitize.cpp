// Host-side run of the FP64 oracle under AddressSanitizer and
// UndefinedBehaviorSanitizer (built and run by tests/test_host_sanitizers.py):
// every entry point of oracle/shrt_oracle.cpp on small grids, square and not,
// through a remap, the reference IC with its rand() noise, the lab-frame
// transform and the diagnostics.  Exits 0 when the runs finish; the
// sanitizers abort on the first error they find.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
struct orc_params {
    int nx, ny;
    double Lx, Ly, delt, cfl_num;
    double nu, kappa, sigma, g, rho0, rho0_prime, rho_eps1;
    int noise, linear;
};
void orc_sizes(int nx, int ny, int* out8);
void orc_fft(int n, int dir, double* data);
void* orc_create(const orc_params* prm);
void orc_destroy(void* h);
void orc_set_threads(void* h, int n);
void orc_initialize_reference(void* h);
void orc_seeded_spectrum(int nx, int ny, double Lx, double Ly, uint64_t seed, double kmin, double kmax, double amp,
                         double expo, double* out);
void orc_set_spectral(void* h, int field, const double* in);
void orc_get_buffer(void* h, int ring, int slot, double* out);
void orc_set_buffer(void* h, int ring, int slot, const double* in);
void orc_get_kdata_bef(void* h, int field, double* out);
void orc_get_physical(void* h, int which, double* out);
void orc_step(void* h, int nsteps);
void orc_set_time(void* h, int istep, double time, double delt);
void orc_get_state(void* h, double* out6);
void orc_get_shift(void* h, double* ky_shift, int* jump);
void orc_set_shift(void* h, const double* ky_shift);
void orc_xtok(void* h, const double* in, double* out);
void orc_ktox(void* h, const double* in, double* out);
void orc_ktox_shear(void* h, const double* in, double* out);
void orc_shearing_field(void* h, const int* jmp, const double* in, double* out);
void orc_poisson_bracket(void* h, const double* in, double* out);
void orc_check_cfl(void* h);
void orc_diagnostics(void* h, double* out3);
}

static int run(int nx, int ny, int threads) {
    int sz[8];
    orc_sizes(nx, ny, sz);
    const int nkx = sz[0], M = sz[6];
    orc_params p{nx, ny, 2 * M_PI, 2 * M_PI, 0.01, 5.0, 1e-2, 1e-2, nx == ny ? 1.0 : 0.0, 1.0, 1.0, 1.0, 1e-2, 1, 0};
    void* h = orc_create(&p);
    orc_set_threads(h, threads);
    orc_initialize_reference(h);
    std::vector<double> spec(2 * (size_t)M), phys((size_t)nx * ny), out(2 * (size_t)M), s(nkx), d3(3), st(6);
    std::vector<int> jmp(nkx);
    orc_seeded_spectrum(nx, ny, 2 * M_PI, 2 * M_PI, 7, 1.0, 6.0, 1e-3, -1.0, spec.data());
    orc_set_spectral(h, 1, spec.data());
    orc_step(h, 120);                      // sigma * t crosses 1 at step 100: one remap on square grids
    orc_get_state(h, st.data());
    for (int f = 0; f < 3; ++f) orc_get_kdata_bef(h, f, out.data());
    for (int r = 0; r < 5; ++r)
        for (int slot = 0; slot < 3; ++slot) orc_get_buffer(h, r, slot, out.data());
    orc_set_buffer(h, 1, 2, out.data());
    orc_get_buffer(h, 0, 0, out.data());
    orc_ktox(h, out.data(), phys.data());
    orc_xtok(h, phys.data(), out.data());
    orc_ktox_shear(h, out.data(), phys.data());
    orc_get_shift(h, s.data(), jmp.data());
    orc_shearing_field(h, jmp.data(), out.data(), spec.data());
    orc_check_cfl(h);
    for (int w = 0; w < 3; ++w) orc_get_physical(h, w, phys.data());
    orc_poisson_bracket(h, out.data(), phys.data());
    orc_diagnostics(h, d3.data());
    orc_set_shift(h, s.data());
    orc_set_time(h, 7, 0.5, 0.005);
    orc_step(h, 3);
    orc_get_state(h, st.data());
    orc_destroy(h);
    std::vector<double> z(2 * 12);
    for (int i = 0; i < 24; ++i) z[i] = std::sin(i);
    orc_fft(12, -1, z.data());            // the direct-DFT path of a length that is not a power of two
    std::printf("oracle %dx%d threads %d: istep %.0f remaps %.0f energy %.6e\n", nx, ny, threads, st[0], st[3], d3[0]);
    return 0;
}

int main() {
    run(32, 32, 1);
    run(16, 32, 1);
    run(32, 32, 2);
    return 0;
}

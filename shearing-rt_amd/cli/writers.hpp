// writers.hpp -- the text outputs of the reference driver (file_access.cu),
// written from host copies of the state in the reference's layouts and
// formats, plus an energy/enstrophy time series.
//
//   output_fields   n%05d_t%09.6f.dat, vel_t%09.6f.dat   (file_access.cu:240-316)
//   output_maxamp   max_amp.dat (appended)                (:320-346)
//   en_spectral     kx_ensp*.dat / ky_ensp*.dat           (:350-447)
//   ks_reim         ksre_t*.dat / ksim_t*.dat             (:451-488)
//   k_data_bef/aft  befk_n*_t*.dat / aftk_n*_t*.dat       (:492-626)
//   output_pbk      pbk.txt / pbk_t%09.6f.dat             (:630-667)
//   energy series   energy.dat (new): time E Z P, FP64 compensated sums
//
// The element type R is the run's precision (float or double); every value
// is printed as the reference prints a cureal, promoted to double.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace shrt_cli {

struct Grid {
    int nx, ny, nkx, nky, nkxh;
    std::vector<double> xx, yy;   // xx[0..nx], yy[0..ny] (time_integral.cu:236-237)
    std::vector<double> kx, ky;   // fourier.cu:123-128
    std::vector<int> kx_index;    // file_access.cu:210-212: kx rows in ascending kx
};

inline FILE* open_or_exit(const std::string& name, const char* mode) {
    FILE* fp = std::fopen(name.c_str(), mode);
    if (!fp) {
        std::fprintf(stderr, "shrt: cannot open %s\n", name.c_str());
        std::exit(1);
    }
    return fp;
}

inline std::string fmt_name(const char* pattern, int n, double t) {
    char buf[256];
    std::snprintf(buf, sizeof buf, pattern, n, t);
    return buf;
}
inline std::string fmt_name_t(const char* pattern, double t) {
    char buf[256];
    std::snprintf(buf, sizeof buf, pattern, t);
    return buf;
}

// Periodic (nx+1) x (ny+1) table of file_access.cu:266-286.
template <typename R, int NF>
void write_table(FILE* fp, const Grid& g, const R* const (&f)[NF]) {
    const int nx = g.nx, ny = g.ny;
    for (int ix = 0; ix <= nx; ++ix) {
        for (int iy = 0; iy <= ny; ++iy) {
            const int sx = ix < nx ? ix : 0, sy = iy < ny ? iy : 0;
            std::fprintf(fp, "%+1.16e %+1.16e", g.xx[ix], g.yy[iy]);
            for (int k = 0; k < NF; ++k) std::fprintf(fp, " %+1.16e", (double)f[k][(size_t)sx * ny + sy]);
            std::fprintf(fp, "\n");
        }
        std::fprintf(fp, "\n");
    }
}

template <typename R>
void output_fields(const Grid& g, int istep, int nwrite, double time, const R* omg, const R* phi, const R* rho,
                   const R* vx, const R* vy) {
    std::printf("time = %g\n", time);
    FILE* fp = open_or_exit(fmt_name("n%05d_t%09.6f.dat", istep / nwrite, time), "w+");
    const R* const f3[3] = {omg, phi, rho};
    write_table<R, 3>(fp, g, f3);
    std::fclose(fp);
    fp = open_or_exit(fmt_name_t("vel_t%09.6f.dat", time), "w+");
    const R* const f2[2] = {vx, vy};
    write_table<R, 2>(fp, g, f2);
    std::fclose(fp);
}

// maxvalue_search (time_integral.cu:396-409): the largest |f|, in R.
template <typename R>
R max_abs(const R* f, size_t n) {
    R m = 0;
    for (size_t i = 0; i < n; ++i) {
        const R a = std::fabs(f[i]);
        if (m < a) m = a;
    }
    return m;
}

template <typename R>
void output_maxamp(double time, R max_omgz, R max_phi, R max_rho) {
    FILE* fp = open_or_exit("max_amp.dat", "a+");
    std::fprintf(fp, "%.10f %+1.16e %+1.16e %+1.16e\n", time, (double)max_omgz, (double)max_phi, (double)max_rho);
    std::fclose(fp);
}

// Spectral fields are interleaved complex f[2*(ikx*nky + iky) + {0,1}].
template <typename R>
void en_spectral(const Grid& g, int istep, double time, const R* ao, const R* ap, const R* ar) {
    if (istep == 0) {
        std::fclose(open_or_exit("kx_ensp.txt", "w+"));
        std::fclose(open_or_exit("ky_ensp.txt", "w+"));
        return;
    }
    const int nkx = g.nkx, nky = g.nky;
    auto sq = [](const R* a, size_t t) {
        const R re = a[2 * t], im = a[2 * t + 1];
        return (R)(re * re + im * im);
    };
    std::vector<R> sx0(nkx), sx1(nkx), sx2(nkx), sy0(nky), sy1(nky), sy2(nky);
    for (int ikx = 0; ikx < nkx; ++ikx)
        for (int iky = 0; iky < nky; ++iky) {
            const size_t t = (size_t)ikx * nky + iky;
            if (iky == 0) { sx0[ikx] = sq(ao, t); sx1[ikx] = sq(ap, t); sx2[ikx] = sq(ar, t); }
            else { sx0[ikx] += sq(ao, t); sx1[ikx] += sq(ap, t); sx2[ikx] += sq(ar, t); }
        }
    FILE* fp = open_or_exit(fmt_name_t("kx_ensp_t%09.6f.dat", time), "w+");
    // the reference labels row ikx (storage order) with kx[kx_index[ikx]]
    for (int ikx = 0; ikx < nkx; ++ikx)
        std::fprintf(fp, "%+1.16e %+1.16e %+1.16e %+1.16e\n", g.kx[g.kx_index[ikx]], (double)(sx0[ikx] / (R)nky),
                     (double)(sx1[ikx] / (R)nky), (double)(sx2[ikx] / (R)nky));
    std::fclose(fp);
    for (int ikx = 0; ikx < nkx; ++ikx)
        for (int iky = 0; iky < nky; ++iky) {
            const size_t t = (size_t)ikx * nky + iky;
            if (ikx == 0) { sy0[iky] = sq(ao, t); sy1[iky] = sq(ap, t); sy2[iky] = sq(ar, t); }
            else { sy0[iky] += sq(ao, t); sy1[iky] += sq(ap, t); sy2[iky] += sq(ar, t); }
        }
    fp = open_or_exit(fmt_name_t("ky_ensp_t%09.6f.dat", time), "w+");
    for (int iky = 0; iky < nky; ++iky)
        std::fprintf(fp, "%+1.16e %+1.16e %+1.16e %+1.16e\n", g.ky[iky], (double)(sy0[iky] / (R)nkx),
                     (double)(sy1[iky] / (R)nkx), (double)(sy2[iky] / (R)nkx));
    std::fclose(fp);
}

template <typename R>
void ks_reim(const Grid& g, double time, const R* ao, const R* ap, const R* ar) {
    for (int part = 0; part < 2; ++part) {
        FILE* fp = open_or_exit(fmt_name_t(part == 0 ? "ksre_t%09.6f.dat" : "ksim_t%09.6f.dat", time), "w+");
        for (int ikx = 0; ikx < g.nkx; ++ikx) {
            for (int iky = 0; iky < g.nky; ++iky) {
                const size_t t = (size_t)g.kx_index[ikx] * g.nky + iky;
                std::fprintf(fp, "%+1.16e %+1.16e %+1.16e %+1.16e %+1.16e\n", g.kx[g.kx_index[ikx]], g.ky[iky],
                             (double)std::fabs(ao[2 * t + part]), (double)std::fabs(ap[2 * t + part]),
                             (double)std::fabs(ar[2 * t + part]));
            }
            std::fprintf(fp, "\n");
        }
        std::fclose(fp);
    }
}

// k_data_bef / k_data_aft body: real then imaginary parts of omega, phi, rho.
template <typename R>
void k_data_dump(const Grid& g, const std::string& name, const R* ao, const R* ap, const R* ar) {
    FILE* fp = open_or_exit(name, "w+");
    const char* titles[3] = {"omg", "phi", "rho"};
    const R* fields[3] = {ao, ap, ar};
    for (int f = 0; f < 3; ++f) {
        std::fprintf(fp, "///////////////////////////// %s //////////////////////////////\n", titles[f]);
        for (int part = 0; part < 2; ++part) {
            for (int ikx = 0; ikx < g.nkx; ++ikx) {
                for (int iky = 0; iky < g.nky; ++iky)
                    std::fprintf(fp, "%+1.16e ", (double)fields[f][2 * ((size_t)ikx * g.nky + iky) + part]);
                std::fprintf(fp, "\n");
            }
            if (part == 0) std::fprintf(fp, "\n\n");
        }
        std::fprintf(fp, f < 2 ? "///////////////////////////////////////////////////////////////\n\n"
                               : "///////////////////////////////////////////////////////////////\n");
    }
    std::fclose(fp);
}

template <typename R>
void k_data_bef(const Grid& g, double time, int istep, int nwrite, const R* ao, const R* ap, const R* ar) {
    k_data_dump(g, fmt_name("befk_n%05d_t%09.6f.dat", istep / nwrite, time), ao, ap, ar);
}

template <typename R>
void k_data_aft(const Grid& g, double time, int istep, int nwrite, const R* ao, const R* ap, const R* ar) {
    if (istep == 0) {
        std::fclose(open_or_exit("after_kdata.txt", "w+"));
        return;
    }
    k_data_dump(g, fmt_name("aftk_n%05d_t%09.6f.dat", istep / nwrite, time), ao, ap, ar);
}

// output_pbk (file_access.cu:630-667): the empty marker at time 0, else the
// advection term in the periodic (nx+1) x (ny+1) table, one value per row.
template <typename R>
void output_pbk(const Grid& g, double time, const R* pb) {
    if (time == 0) {
        std::fclose(open_or_exit("pbk.txt", "w+"));
        return;
    }
    FILE* fp = open_or_exit(fmt_name_t("pbk_t%09.6f.dat", time), "w+");
    const R* const f1[1] = {pb};
    write_table<R, 1>(fp, g, f1);
    std::fclose(fp);
}

inline void energy_series(bool first, double time, double E, double Z, double P) {
    FILE* fp = open_or_exit("energy.dat", first ? "w+" : "a+");
    std::fprintf(fp, "%.10f %+1.16e %+1.16e %+1.16e\n", time, E, Z, P);
    std::fclose(fp);
}

}  // namespace shrt_cli

// mmc_main.cpp -- command-line annealer, the drop-in for the reference's
// executable (src/main.cu:90-1266):
//
//   mmc_run input.ini [-gpu=N] [-tensorcache="dir/"] [-debugfields]
//
// Reads the reference's ini keys, writes <ini-dir>/<ini-stem>.output/ with
// energy.txt ("%f %ld %g": T, sweep, E[J]) and OVF snapshots
// "%06ld_%.2fK_%ld.ovf" (counter, T, sweep), loads/stores the tensor cache,
// and runs the temperature-stage loop with Ctrl-C handling.  -debugfields
// also writes, at every snapshot, the per-term fields B_demag, B_anis,
// B_exch, B_dmi, B_zeeman, B_effective [T] and E_density [J/m^3, in all three
// components] as OVF files of those names (overwritten), as the reference
// does under its compile-time DEBUG_H_FIELD switch (src/main.cu:965-1064).
#include "mmc.h"
#include "mmc_io.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <string>
#include <vector>

namespace fs = std::filesystem;

static volatile sig_atomic_t g_stop = 0;          // src/main.cu:82-88
static void on_sigint(int) { g_stop = 1; }

static void usage()
{
    std::printf("need input file. \n");
    std::printf("arg format:  input_cfg_file.ini  -gpu=0   -tensorcache=\"/mnt/cachepath/\" "
                "[-debugfields]\n");
}

static void echo_config(const mmc_config &c)   // cout_input_ini, src/parse_input_ini.h:49-114
{
    std::printf("====================display ini file============\n");
    std::printf("nx: %d\nny: %d\nnz: %d\n", c.nx, c.ny, c.nz);
    std::printf("dx: %g\ndy: %g\ndz: %g\n", c.dx, c.dy, c.dz);
    std::printf("Bextx: %g\nBexty: %g\nBextz: %g\n", c.Bextx, c.Bexty, c.Bextz);
    std::printf("pbc_x: %d\npbc_y: %d\npbc_z: %d\n", c.pbc_x, c.pbc_y, c.pbc_z);
    std::printf("ms: %g\nAex: %g\nDind: %g\n", c.ms, c.Aex, c.Dind);
    std::printf("anisUx: %g\nanisUy: %g\nanisUz: %g\nKu1: %g\n", c.anisUx, c.anisUy, c.anisUz, c.Ku1);
    std::printf("use_random_init: %d\nrandomseed: %lld\novf_filename: %s\n", c.use_random_init,
                c.randomseed, c.ovf_filename);
    std::printf("Temperature_start: %g\nTemperature_end: %g\nTemperature_step: %g\n",
                c.Temperature_start, c.Temperature_end, c.Temperature_step);
    std::printf("Temperature_use_exp: %d\n", c.Temperature_use_exp);
    std::printf("circle_per_stage: %ld\nterminal_output_period: %ld\n", c.circle_per_stage,
                c.terminal_output_period);
    std::printf("energy_output_period: %ld\nms_output_period: %ld\n", c.energy_output_period,
                c.ms_output_period);
    std::printf("rand_block_size: %d\ncal_demag_flag: %d\n", c.rand_block_size, c.cal_demag_flag);
    std::printf("precision: %s\n", c.precision == MMC_FP64 ? "fp64" : "fp32");
    std::printf("---------------------END display ini file------------\n");
}

static bool due(long s, long period) { return period > 0 && s % period == 0; }

// Per-term fields [T] and the energy density of instance 0 as OVF files in
// `dir` (the reference's DEBUG_H_FIELD dumps, src/main.cu:965-1064).
static int write_debug_fields(mmc_ctx *ctx, const fs::path &dir, const mmc_ovf_info &info,
                              long N, int fmt)
{
    const double mu0 = 4e-7 * 3.14159265358;          // src/main.cu:244
    std::vector<double> hx(N), hy(N), hz(N);
    static const struct { const char *name; int terms; } f[] = {
        {"B_demag.ovf", MMC_DEMAG}, {"B_anis.ovf", MMC_ANIS}, {"B_exch.ovf", MMC_EXCH},
        {"B_dmi.ovf", MMC_DMI}, {"B_zeeman.ovf", MMC_ZEEMAN}, {"B_effective.ovf", MMC_ALL}};
    for (const auto &t : f) {
        int rc = mmc_field(ctx, 0, t.terms, hx.data(), hy.data(), hz.data());
        if (rc != MMC_OK) return rc;
        for (long i = 0; i < N; ++i) { hx[i] *= mu0; hy[i] *= mu0; hz[i] *= mu0; }
        if (mmc_ovf_write((dir / t.name).c_str(), &info, hx.data(), hy.data(), hz.data(), fmt,
                          "B") != MMC_OK)
            std::fprintf(stderr, "cannot write %s\n", t.name);
    }
    int rc = mmc_energy_density(ctx, 0, MMC_ALL, hx.data());
    if (rc != MMC_OK) return rc;
    if (mmc_ovf_write((dir / "E_density.ovf").c_str(), &info, hx.data(), hx.data(), hx.data(), fmt,
                      "E_density") != MMC_OK)
        std::fprintf(stderr, "cannot write E_density.ovf\n");
    return MMC_OK;
}

// next sweep index >= s+1 at which any output falls due (or `end`)
static long next_due(long s, long end, const long *periods, int n)
{
    long nx = end;
    for (int i = 0; i < n; ++i) {
        long p = periods[i];
        if (p <= 0) continue;
        long k = ((s + 1 + p - 1) / p) * p;
        if (k < nx) nx = k;
    }
    return nx;
}

int main(int argc, char **argv)
{
    setbuf(stdout, nullptr);
    if (argc == 1) { usage(); return 0; }
    std::string ini, cache_dir;
    int gpu = -1;
    bool debug_fields = false;
    for (int i = 1; i < argc; ++i) {                 // src/GPU_selector.h:11-68
        std::string a = argv[i];
        if (a.rfind("-gpu=", 0) == 0) gpu = std::atoi(a.c_str() + 5);
        else if (a.rfind("-tensorcache=", 0) == 0) {
            cache_dir = a.substr(13);
            if (cache_dir.size() >= 2 && cache_dir.front() == '"' && cache_dir.back() == '"')
                cache_dir = cache_dir.substr(1, cache_dir.size() - 2);
        } else if (a == "-debugfields") debug_fields = true;
        else if (!a.empty() && a[0] != '-' && ini.empty()) ini = a;
    }
    if (ini.empty()) { usage(); return 0; }
    if (gpu >= 0 && hipSetDevice(gpu) != hipSuccess) {
        std::fprintf(stderr, "cannot select GPU %d\n", gpu);
        return 1;
    }
    std::signal(SIGINT, on_sigint);

    fs::path ini_path = fs::absolute(fs::path(ini));
    std::printf("input_cfg_path is : %s\n", ini_path.c_str());
    mmc_config cfg;
    char msg[1024];
    if (mmc_ini_parse(ini_path.c_str(), &cfg, msg, sizeof msg) != MMC_OK) {
        std::fprintf(stderr, "ini error: %s\n", msg);
        return 1;
    }
    echo_config(cfg);

    // output folder, src/main.cu:193-239
    fs::path out_dir = ini_path.parent_path() / (ini_path.stem().string() + ".output");
    std::error_code ec;
    if (!fs::exists(out_dir)) {
        if (!fs::create_directories(out_dir, ec)) {
            std::printf("Fail to create output folder: %s\n", out_dir.c_str());
            return 1;
        }
        std::printf("Successfully create output folder: %s\n", out_dir.c_str());
    } else {
        for (auto &e : fs::directory_iterator(out_dir)) {
            std::string ext = e.path().extension().string();
            if (ext == ".ovf" || ext == ".png" || ext == ".jpg" || ext == ".jpeg") fs::remove(e.path(), ec);
        }
    }
    std::printf("output_path: %s\n", out_dir.c_str());
    const std::string energy_file = (out_dir / "energy.txt").string();
    if (FILE *fp = std::fopen(energy_file.c_str(), "w")) std::fclose(fp);

    long long seed = cfg.randomseed;
    if (seed < 0) {                                   // src/main.cu:246-250
        seed = (long long)std::time(nullptr);
        std::printf("Use random seed: %lld\n", seed);
    }

    mmc_params p{};
    p.nx = cfg.nx; p.ny = cfg.ny; p.nz = cfg.nz;
    p.dx = cfg.dx; p.dy = cfg.dy; p.dz = cfg.dz;
    p.pbc_x = cfg.pbc_x; p.pbc_y = cfg.pbc_y; p.pbc_z = cfg.pbc_z;
    p.ms = cfg.ms; p.aex = cfg.Aex; p.dind = cfg.Dind;
    p.anis_ux = cfg.anisUx; p.anis_uy = cfg.anisUy; p.anis_uz = cfg.anisUz; p.ku1 = cfg.Ku1;
    p.bext_x = cfg.Bextx; p.bext_y = cfg.Bexty; p.bext_z = cfg.Bextz;
    p.cal_demag = cfg.cal_demag_flag;
    p.seed = (unsigned long long)seed;
    mmc_ctx *ctx = nullptr;
    if (mmc_create(&p, cfg.precision, &ctx) != MMC_OK) {
        std::fprintf(stderr, "mmc_create: %s\n", mmc_last_error(ctx));
        mmc_destroy(ctx);
        return 1;
    }
    int dims[6];
    mmc_dims(ctx, 0, dims);
    const long N = (long)cfg.nx * cfg.ny * cfg.nz;
    std::printf("\nDimension: nx ny nz meshsize\n%d %d %d %ld\n", dims[0], dims[1], dims[2], N);
    std::printf("nx_padded ny_padded nz_padded meshsize_padded\n%d %d %d %ld\n", dims[3], dims[4],
                dims[5], (long)dims[3] * dims[4] * dims[5]);

    mmc_ovf_info info{cfg.nx, cfg.ny, cfg.nz, cfg.dx, cfg.dy, cfg.dz};
    std::vector<double> mx(N), my(N), mz(N);
    int rc = MMC_OK;
    if (cfg.use_random_init) {
        rc = mmc_init_random_m(ctx, 0);
    } else {                                          // src/main.cu:317-323
        if (!fs::exists(cfg.ovf_filename)) {
            std::printf("OVF file not found: %s\n", cfg.ovf_filename);
            mmc_destroy(ctx);
            return 1;
        }
        mmc_ovf_info fin;
        rc = mmc_ovf_read(cfg.ovf_filename, &fin, mx.data(), my.data(), mz.data(), N, msg, sizeof msg);
        if (rc == MMC_OK && !mmc_ovf_mesh_equal(&fin, &info)) {
            std::snprintf(msg, sizeof msg, "mesh of %s does not match the ini", cfg.ovf_filename);
            rc = MMC_EINVAL;
        }
        if (rc != MMC_OK) {
            std::printf("Error, quit now. (%s)\n", msg);
            mmc_destroy(ctx);
            return 1;
        }
        for (long i = 0; i < N; ++i) {           // normalise (a resumed state may drift)
            double l = std::sqrt(mx[i] * mx[i] + my[i] * my[i] + mz[i] * mz[i]);
            if (l > 0) { mx[i] /= l; my[i] /= l; mz[i] /= l; } else { mx[i] = my[i] = 0; mz[i] = 1; }
        }
        rc = mmc_set_m(ctx, 0, mx.data(), my.data(), mz.data());
    }
    if (rc != MMC_OK) { std::fprintf(stderr, "%s\n", mmc_last_error(ctx)); mmc_destroy(ctx); return 1; }

    std::printf("Calculating demag tensor...\n");
    auto t0 = std::chrono::steady_clock::now();
    rc = mmc_build_tensor(ctx, cache_dir.empty() ? nullptr : cache_dir.c_str());
    if (rc != MMC_OK) { std::fprintf(stderr, "%s\n", mmc_last_error(ctx)); mmc_destroy(ctx); return 1; }
    std::printf("Done (%.3f s).\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    // replay captured sweeps and keep the fastest sweep schedule (every
    // schedule gives the same Markov chain)
    if ((rc = mmc_set_option(ctx, "graph", 1)) == MMC_OK) {
        char rep[2048];
        rc = mmc_autotune(ctx, 10, rep, sizeof rep);
        if (rc == MMC_OK) std::printf("%s", rep);
    }
    if (rc != MMC_OK) { std::fprintf(stderr, "%s\n", mmc_last_error(ctx)); mmc_destroy(ctx); return 1; }
    std::printf("Start...\n");

    // temperature-stage loop, src/main.cu:801-1139
    double T = cfg.Temperature_start;
    long stage_idx = 0, file_count = 0;
    const long periods[3] = {cfg.terminal_output_period, cfg.energy_output_period, cfg.ms_output_period};
    while (T > cfg.Temperature_end && !g_stop) {
        std::printf("\n\nTemperature: %.2f K\n", T);
        if ((rc = mmc_set_temperature(ctx, -1, T)) != MMC_OK) break;
        auto ts = std::chrono::steady_clock::now();
        long s = 0;
        while (s < cfg.circle_per_stage && !g_stop) {
            if (due(s, cfg.terminal_output_period)) {
                std::printf("Stage: %ld", s);
                if (s > 0) {
                    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
                    std::printf(", elapsed time: %f s, speed: %f stages/s ", el, s / el);
                } else {
                    std::printf(" ");
                }
            }
            if (due(s, cfg.ms_output_period)) {
                if ((rc = mmc_get_m(ctx, 0, mx.data(), my.data(), mz.data())) != MMC_OK) break;
                char name[200];
                std::snprintf(name, sizeof name, "%06ld_%.2fK_%ld.ovf", file_count, T, s);
                if (mmc_ovf_write((out_dir / name).c_str(), &info, mx.data(), my.data(), mz.data(),
                                  cfg.ovf_format, "m") != MMC_OK)
                    std::fprintf(stderr, "cannot write %s\n", name);
                ++file_count;
                if (debug_fields &&
                    (rc = write_debug_fields(ctx, out_dir, info, N, cfg.ovf_format)) != MMC_OK)
                    break;
            }
            if (due(s, cfg.energy_output_period)) {
                double E = 0;
                if ((rc = mmc_energy(ctx, 0, &E, nullptr, nullptr)) != MMC_OK) break;
                std::printf("\nEnergy:  %g\n", E);
                if (FILE *fp = std::fopen(energy_file.c_str(), "a")) {
                    std::fprintf(fp, "%f %ld %g\n", T, s, E);
                    std::fclose(fp);
                }
            }
            long e = next_due(s, cfg.circle_per_stage, periods, 3);
            if (e - s > 1000) e = s + 1000;          // poll Ctrl-C at least every 1000 sweeps
            if ((rc = mmc_sweeps(ctx, (uint32_t)stage_idx, (uint32_t)s, (uint32_t)(e - s))) != MMC_OK) break;
            if ((rc = mmc_sync(ctx)) != MMC_OK) break;
            s = e;
        }
        if (rc != MMC_OK) break;
        ++stage_idx;
        if (g_stop) break;
        T = cfg.Temperature_use_exp ? T * cfg.Temperature_step : T - cfg.Temperature_step;
    }
    if (rc != MMC_OK) {
        std::fprintf(stderr, "error: %s\n", mmc_last_error(ctx));
        mmc_destroy(ctx);
        return 1;
    }
    if (g_stop) std::printf("\nInterrupted.\n");
    mmc_destroy(ctx);
    std::printf("~~~~~~~~~~~~~~END~~~~~~~~~~~~~~~~~~~~~~~~~ \n");
    return 0;
}

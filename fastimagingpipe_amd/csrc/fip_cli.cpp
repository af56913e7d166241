// sharedlibrary_gpu: command-line drop-in for the reference's
// pipe_interface.cpp:139-189 over the fip C ABI.
//
//   sharedlibrary_gpu Visreal Visimag Bin Vin Image_Size Number_of_Baselines
//                     Cell_Size Number_of_Snapshots Tile_Size Output.fits
//   sharedlibrary_gpu ... Number_of_Baselines Frequency Cell_Size ...   (README.md:30 form;
//                     Frequency is accepted and ignored: B is already in wavelengths)
//   sharedlibrary_gpu --combined input.fits Image_Size Cell_Size Tile_Size Output.fits
//                     [--all-windows]
//       3-extension file of the preprocessor (SVDPreprocessMain.cpp:181-262):
//       TRANSFORMATION MATRICES (3,3,S), VISIBILITIES (2,N,S), R-COORDINATES (2,N,S).
//       Default: snapshots 0..2 -> one tLISI matrix; --all-windows: one matrix per
//       window (s, s+1, s+2) written as a (I/T, I/T, S-2) cube.
//   Either form may end with  --tau X : list the triggered tiles, tLISI < X
//   (the reference computes tLISI but defines no threshold).
#include "fip.h"

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

void usage()
{
    std::fprintf(stderr,
                 "usage: sharedlibrary_gpu Visreal.fits Visimag.fits Bin.fits Vin.fits Image_Size\n"
                 "                         Number_of_Baselines [Frequency] Cell_Size Number_of_Snapshots\n"
                 "                         Tile_Size Output.fits\n"
                 "       sharedlibrary_gpu --combined input.fits Image_Size Cell_Size Tile_Size Output.fits\n"
                 "                         [--all-windows]\n"
                 "       either form may end with --tau X (list tiles with tLISI < X)\n");
}

bool parse_size(const char* s, size_t* out)
{
    if (!s || !*s || *s == '-') return false;
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(s, &end, 10);
    if (errno || *end) return false;
    *out = (size_t)v;
    return true;
}

bool parse_float(const char* s, float* out)
{
    if (!s || !*s) return false;
    char* end = nullptr;
    errno = 0;
    const float v = std::strtof(s, &end);
    if (errno || *end) return false;
    *out = v;
    return true;
}

// --tau X, stripped from the argument list by main (negative: not given).
float g_tau = -1.0f;

void report_triggers(const float* t, size_t un, long window)
{
    if (g_tau < 0.0f) return;
    size_t n = 0;
    for (size_t q = 0; q < un * un; ++q) n += t[q] < g_tau;
    if (window >= 0)
        std::printf("window %ld: %zu tile(s) with tLISI < %g\n", window, n, (double)g_tau);
    else
        std::printf("%zu tile(s) with tLISI < %g\n", n, (double)g_tau);
    for (size_t q = 0; q < un * un; ++q)
        if (t[q] < g_tau) std::printf("  tile row %zu col %zu tLISI %.9g\n", q / un, q % un, (double)t[q]);
}

struct Fits {
    float* data = nullptr;
    long dims[3] = {0, 0, 0};
    ~Fits() { fip_free(data); }
    size_t count() const { return (size_t)dims[0] * dims[1] * dims[2]; }
};

// pipe_interface.cpp:12-96: primary image, NAXIS 1 or 2, converted to float.
bool read_input(const char* path, Fits& f)
{
    if (fip_fits_read(path, &f.data, f.dims) != FIP_OK) {
        std::fprintf(stderr, "ERROR: %s\n", fip_last_error());
        return false;
    }
    const int naxis = f.dims[0] == 0 ? 0 : (f.dims[2] > 1 ? 3 : 2);
    if (naxis == 0 || naxis > 2) {
        std::fprintf(stderr, "ERROR: Unsupported naxis for file %s\n", path);
        return false;
    }
    return true;
}

int run_separate(int argc, char** argv)
{
    const bool readme_form = (argc == 12);
    const int o = readme_form ? 1 : 0;
    size_t image = 0, nbl = 0, nsnap = 0, tile = 0;
    float cell = 0.0f, freq = 0.0f;
    if (!parse_size(argv[5], &image) || !parse_size(argv[6], &nbl) ||
        (readme_form && !parse_float(argv[7], &freq)) || !parse_float(argv[7 + o], &cell) ||
        !parse_size(argv[8 + o], &nsnap) || !parse_size(argv[9 + o], &tile)) {
        std::fprintf(stderr, "ERROR: malformed numeric argument\n");
        usage();
        return 1;
    }
    const char* out = argv[10 + o];
    Fits vr, vi, b, v;
    if (!read_input(argv[1], vr) || !read_input(argv[2], vi) || !read_input(argv[3], b) || !read_input(argv[4], v))
        return 1;
    // The reference reads these arrays without checking their sizes; here a
    // mismatch is an error instead of an out-of-bounds read.
    if (vr.count() < nbl * nsnap || vi.count() < nbl * nsnap || b.count() < 2 * nbl * nsnap ||
        v.count() < 9 * nsnap) {
        std::fprintf(stderr,
                     "ERROR: input sizes (%zu, %zu, %zu, %zu) do not hold %zu snapshots of %zu baselines\n",
                     vr.count(), vi.count(), b.count(), v.count(), nsnap, nbl);
        return 1;
    }
    if (tile == 0 || tile > image) {
        std::fprintf(stderr, "ERROR: Tile_Size must be in [1, Image_Size]\n");
        return 1;
    }
    const size_t un = image / tile;
    std::vector<float> result(un * un);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = fip_pipe(vr.data, vi.data, b.data, v.data, result.data(), nbl, image, nsnap, cell, tile);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != FIP_OK && rc != FIP_ERR_DEGENERATE) {
        std::fprintf(stderr, "ERROR: %s: %s\n", fip_status_string(rc), fip_last_error());
        return 1;
    }
    if (rc == FIP_ERR_DEGENERATE) std::fprintf(stderr, "WARNING: %s\n", fip_last_error());
    std::printf("Time elapsed: %.3f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
    report_triggers(result.data(), un, -1);
    if (fip_fits_write(out, result.data(), (long)un, (long)un) != FIP_OK) {
        std::fprintf(stderr, "Error writing FITS image: %s\n", fip_last_error());
        return 1;
    }
    return 0;
}

int run_combined(int argc, char** argv)
{
    // --combined input.fits Image_Size Cell_Size Tile_Size Output.fits [--all-windows]
    if (argc != 7 && argc != 8) {
        usage();
        return 1;
    }
    const bool all = (argc == 8);
    if (all && std::strcmp(argv[7], "--all-windows") != 0) {
        usage();
        return 1;
    }
    size_t image = 0, tile = 0;
    float cell = 0.0f;
    if (!parse_size(argv[3], &image) || !parse_float(argv[4], &cell) || !parse_size(argv[5], &tile)) {
        std::fprintf(stderr, "ERROR: malformed numeric argument\n");
        return 1;
    }
    Fits v, vis, r;
    if (fip_fits_read_ext(argv[2], "TRANSFORMATION MATRICES", &v.data, v.dims) != FIP_OK ||
        fip_fits_read_ext(argv[2], "VISIBILITIES", &vis.data, vis.dims) != FIP_OK ||
        fip_fits_read_ext(argv[2], "R-COORDINATES", &r.data, r.dims) != FIP_OK) {
        std::fprintf(stderr, "ERROR: %s\n", fip_last_error());
        return 1;
    }
    const size_t S = (size_t)v.dims[2], N = (size_t)vis.dims[1];
    if (v.dims[0] != 3 || v.dims[1] != 3 || vis.dims[0] != 2 || r.dims[0] != 2 || (size_t)r.dims[1] != N ||
        (size_t)vis.dims[2] != S || (size_t)r.dims[2] != S) {
        std::fprintf(stderr, "ERROR: combined FITS extensions have inconsistent shapes\n");
        return 1;
    }
    if (tile == 0 || tile > image) {
        std::fprintf(stderr, "ERROR: Tile_Size must be in [1, Image_Size]\n");
        return 1;
    }
    std::vector<float> vr(S * N), vi(S * N);
    for (size_t q = 0; q < S * N; ++q) {
        vr[q] = vis.data[2 * q];
        vi[q] = vis.data[2 * q + 1];
    }
    const size_t un = image / tile;
    int rc;
    if (!all) {
        std::vector<float> result(un * un);
        rc = fip_pipe(vr.data(), vi.data(), r.data, v.data, result.data(), N, image, S, cell, tile);
        if (rc == FIP_OK) report_triggers(result.data(), un, -1);
        if (rc == FIP_OK || rc == FIP_ERR_DEGENERATE)
            rc = fip_fits_write(argv[6], result.data(), (long)un, (long)un) == FIP_OK ? rc : FIP_ERR_IO;
    } else {
        if (S < 3) {
            std::fprintf(stderr, "ERROR: at least 3 snapshots are needed\n");
            return 1;
        }
        // Zero-padded snapshots (SVDPreprocessMain.cpp:264-291): every snapshot
        // holds N samples and is normalised by N, as the reference would.
        std::vector<int> nvis(S, (int)N);
        std::vector<float> result((S - 2) * un * un);
        rc = fip_batch_host(vr.data(), vi.data(), FIP_VIS_F32, r.data, v.data, nvis.data(), N, S, N, image, cell,
                            tile, result.data());
        if (rc == FIP_OK)
            for (size_t w = 0; w + 2 < S; ++w) report_triggers(result.data() + w * un * un, un, (long)w);
        if (rc == FIP_OK) rc = fip_fits_write3(argv[6], result.data(), (long)un, (long)un, (long)(S - 2));
    }
    if (rc != FIP_OK && rc != FIP_ERR_DEGENERATE) {
        std::fprintf(stderr, "ERROR: %s: %s\n", fip_status_string(rc), fip_last_error());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc >= 3 && std::strcmp(argv[argc - 2], "--tau") == 0) {
        if (!parse_float(argv[argc - 1], &g_tau) || !(g_tau >= 0.0f)) {
            std::fprintf(stderr, "ERROR: --tau needs a non-negative number\n");
            return 1;
        }
        argc -= 2;
    }
    if (argc >= 2 && std::strcmp(argv[1], "--combined") == 0) return run_combined(argc, argv);
    if (argc != 11 && argc != 12) {
        usage();
        return 1;
    }
    return run_separate(argc, argv);
}

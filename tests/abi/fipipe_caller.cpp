// A C++ caller of the reference's FIpipe entry point, compiled against the
// reference's own prototype (pipe_interface.cpp:8) and linked with -lfip, as
// an unmodified pipe_interface.cpp would be.  It does not include fip.h.
//
//   fipipe_caller <in.bin> <out.bin> <num_baselines> <image_size> <num_snapshots> <cell_size> <unit_size>
//
// in.bin holds Visreal, Visimag, Bin, Vin back to back as float32
// (num_snapshots*N, num_snapshots*N, num_snapshots*N*2, num_snapshots*9);
// out.bin receives the (image/unit)^2 float32 tLISI values.  Exit status is
// FIpipe's return value (0 on success).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int FIpipe(float* Visreal, float* Visimag, float* Bin, float* Vin, float* result_array, size_t num_baselines,
           size_t image_size, size_t num_snapshots, float cell_size, size_t unit_size);

int main(int argc, char** argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s in.bin out.bin N I nsnap cell T\n", argv[0]);
        return 2;
    }
    const size_t N = std::stoul(argv[3]), I = std::stoul(argv[4]), S = std::stoul(argv[5]);
    const float cell = std::stof(argv[6]);
    const size_t T = std::stoul(argv[7]);
    std::vector<float> vr(S * N), vi(S * N), b(S * N * 2), v(S * 9), out((I / T) * (I / T));
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    bool ok = std::fread(vr.data(), 4, vr.size(), f) == vr.size() && std::fread(vi.data(), 4, vi.size(), f) == vi.size() &&
              std::fread(b.data(), 4, b.size(), f) == b.size() && std::fread(v.data(), 4, v.size(), f) == v.size();
    std::fclose(f);
    if (!ok) return 3;
    const int rc = FIpipe(vr.data(), vi.data(), b.data(), v.data(), out.data(), N, I, S, cell, T);
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 3;
    ok = std::fwrite(out.data(), 4, out.size(), g) == out.size();
    std::fclose(g);
    return ok ? rc : 3;
}

// Synthetic FIP-TOI workloads: array layout, earth-rotation uvw, FP64 DFT
// visibilities of a power-law point-source sky with transients, Gaussian
// noise, optional per-snapshot flagging, and the SVD-plane projection of each
// snapshot in FP64 (SVD_MFS.py:38-110 conventions).
#include "fip.h"

#include "fip_context.hpp"
#include "fip_eigen.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kC = 299792458.0;                     // SVD_MFS.py:7
constexpr double kOmegaEarth = 2.0 * M_PI / 86164.0905;   // sidereal rad/s

// xoshiro256** seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        for (auto& w : s) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

template <class F>
void parallel_for(size_t n, F&& f)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int fip_synth(int nant, double radius_m, size_t nsnap, double dt_s, double freq_hz, double dec_deg,
                         double lat_deg, size_t image_size, float cell, int nsrc, double smin, double smax,
                         double index, double noise_sigma, int ntransient, const int* transient_snap,
                         double transient_flux, double flag_frac_max, unsigned long long seed, float* vis_re,
                         float* vis_im, float* bin, float* vin, int* nvis_out)
{
    using fip::set_error;
    try {
        if (nant < 2 || nsnap == 0 || !vis_re || !vis_im || !bin || !vin)
            return set_error(FIP_ERR_ARG, "fip_synth: bad arguments");
        if (ntransient > 0 && !transient_snap) return set_error(FIP_ERR_ARG, "fip_synth: transient_snap missing");
        if (!(radius_m > 0) || !(freq_hz > 0) || !(smin > 0) || !(smax >= smin))
            return set_error(FIP_ERR_ARG, "fip_synth: bad physical parameters");
        Rng rng(seed);
        // Antennas: log-uniform radius in [R/100, R], uniform azimuth, on the ground plane.
        std::vector<double> E(nant), N(nant);
        for (int a = 0; a < nant; ++a) {
            const double r = radius_m * std::pow(10.0, -2.0 * (1.0 - rng.uniform()));
            const double az = 2.0 * M_PI * rng.uniform();
            E[a] = r * std::cos(az);
            N[a] = r * std::sin(az);
        }
        const size_t nbl = (size_t)nant * (nant - 1) / 2;
        // Baselines in equatorial XYZ (local ENU with U = 0, site latitude lat).
        const double lat = lat_deg * M_PI / 180.0, dec = dec_deg * M_PI / 180.0;
        std::vector<double> X(nbl), Y(nbl), Z(nbl);
        {
            size_t k = 0;
            for (int i = 0; i < nant; ++i)
                for (int j = i + 1; j < nant; ++j, ++k) {
                    const double dE = E[j] - E[i], dN = N[j] - N[i];
                    X[k] = -std::sin(lat) * dN;
                    Y[k] = dE;
                    Z[k] = std::cos(lat) * dN;
                }
        }
        // Sky: power-law fluxes dN/dS ~ S^index within the central field.
        const double fov = 0.4 * (double)image_size * (double)cell;
        struct Src { double l, m, n1, s; int snap; };
        std::vector<Src> sky;
        const double a1 = index + 1.0;
        for (int k = 0; k < nsrc; ++k) {
            Src s;
            s.l = fov * (2.0 * rng.uniform() - 1.0);
            s.m = fov * (2.0 * rng.uniform() - 1.0);
            const double u = rng.uniform();
            s.s = std::fabs(a1) < 1e-12 ? smin * std::pow(smax / smin, u)
                                        : std::pow(std::pow(smin, a1) + u * (std::pow(smax, a1) - std::pow(smin, a1)), 1.0 / a1);
            s.n1 = std::sqrt(std::max(0.0, 1.0 - s.l * s.l - s.m * s.m)) - 1.0;
            s.snap = -1;
            sky.push_back(s);
        }
        for (int k = 0; k < ntransient; ++k) {
            Src s;
            s.l = 0.75 * fov * (2.0 * rng.uniform() - 1.0);
            s.m = 0.75 * fov * (2.0 * rng.uniform() - 1.0);
            s.s = transient_flux;
            s.n1 = std::sqrt(std::max(0.0, 1.0 - s.l * s.l - s.m * s.m)) - 1.0;
            s.snap = transient_snap[k];
            sky.push_back(s);
        }
        const double lam_scale = freq_hz / kC;
        std::vector<double> u(nbl), v(nbl), w(nbl);
        std::vector<std::complex<double>> vis(nbl);
        std::vector<uint8_t> flagged(nbl);
        std::vector<size_t> idx(nbl);
        for (size_t s = 0; s < nsnap; ++s) {
            const double H = (double)s * dt_s * kOmegaEarth;
            const double sH = std::sin(H), cH = std::cos(H), sd = std::sin(dec), cd = std::cos(dec);
            for (size_t k = 0; k < nbl; ++k) {
                u[k] = (sH * X[k] + cH * Y[k]) * lam_scale;
                v[k] = (-sd * cH * X[k] + sd * sH * Y[k] + cd * Z[k]) * lam_scale;
                w[k] = (cd * cH * X[k] - cd * sH * Y[k] + sd * Z[k]) * lam_scale;
            }
            // FP64 DFT, V = sum S exp(-2 pi i (u l + v m + w (n - 1)))
            parallel_for(nbl, [&](size_t b, size_t e) {
                for (size_t k = b; k < e; ++k) {
                    double re = 0.0, im = 0.0;
                    for (const Src& q : sky) {
                        if (q.snap >= 0 && (size_t)q.snap != s) continue;
                        const double ph = -2.0 * M_PI * (u[k] * q.l + v[k] * q.m + w[k] * q.n1);
                        re += q.s * std::cos(ph);
                        im += q.s * std::sin(ph);
                    }
                    vis[k] = {re, im};
                }
            });
            for (size_t k = 0; k < nbl; ++k) vis[k] += std::complex<double>(noise_sigma * rng.normal(), noise_sigma * rng.normal());
            // Projection of this snapshot's uvw (PCA over all rows, flagged included).
            double mu[3] = {0, 0, 0};
            for (size_t k = 0; k < nbl; ++k) {
                mu[0] += u[k];
                mu[1] += v[k];
                mu[2] += w[k];
            }
            for (double& m : mu) m /= (double)nbl;
            double cov[3][3] = {{0}};
            for (size_t k = 0; k < nbl; ++k) {
                const double d[3] = {u[k] - mu[0], v[k] - mu[1], w[k] - mu[2]};
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) cov[i][j] += d[i] * d[j];
            }
            for (auto& row : cov)
                for (double& c : row) c /= (double)(nbl - 1);
            double V[3][3];
            fip::pca_basis(cov, V);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) vin[s * 9 + 3 * i + j] = (float)V[i][j];
            // Flagging: a random fraction in [0, flag_frac_max) of this snapshot.
            std::fill(flagged.begin(), flagged.end(), 0);
            if (flag_frac_max > 0.0) {
                const size_t nflag = (size_t)std::floor(rng.uniform() * flag_frac_max * (double)nbl);
                for (size_t k = 0; k < nbl; ++k) idx[k] = k;
                for (size_t k = 0; k < nflag; ++k) {
                    const size_t j = k + (size_t)(rng.uniform() * (double)(nbl - k));
                    std::swap(idx[k], idx[std::min(j, nbl - 1)]);
                    flagged[idx[k]] = 1;
                }
            }
            size_t out = 0;
            for (size_t k = 0; k < nbl; ++k) {
                if (flagged[k]) continue;
                const size_t q = s * nbl + out;
                vis_re[q] = (float)vis[k].real();
                vis_im[q] = (float)vis[k].imag();
                // B = D . [V0 V1], un-centred (SVD_MFS.py:105-110)
                bin[2 * q] = (float)(u[k] * V[0][0] + v[k] * V[0][1] + w[k] * V[0][2]);
                bin[2 * q + 1] = (float)(u[k] * V[1][0] + v[k] * V[1][1] + w[k] * V[1][2]);
                ++out;
            }
            for (size_t k = out; k < nbl; ++k) {   // zero padding (SVDPreprocessMain.cpp:264-291)
                const size_t q = s * nbl + k;
                vis_re[q] = vis_im[q] = 0.0f;
                bin[2 * q] = bin[2 * q + 1] = 0.0f;
            }
            if (nvis_out) nvis_out[s] = (int)out;
        }
        return FIP_OK;
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "fip_synth: host allocation failed");
    }
}

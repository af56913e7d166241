// Host-side plan math (see fip_plan.hpp).  All tables are computed in FP64
// and rounded once to FP32.
#include "fip_plan.hpp"

#include "fip_internal.hpp"

#include <cmath>

namespace fip {

size_t grid_size_for(size_t image)
{
    const float num = static_cast<float>(1.5 * static_cast<double>(static_cast<float>(image)));
    return static_cast<size_t>(std::ceil(num));
}

bool image_size_supported(size_t image)
{
    if (image < 64 || image % 32 != 0 || image > 8192) return false;
    const size_t g = grid_size_for(image);
    return fft_size_supported(static_cast<int>(g)) && fft_size_supported(static_cast<int>(g / 2));
}

double es_kernel(double x)
{
    const double xx = x * x;
    return xx > 1.0 ? 0.0 : std::exp(kBeta * (std::sqrt(1.0 - xx) - 1.0));
}

void gauss_legendre28_half(double nodes[14], double weights[14])
{
    // Newton iteration on P_28 from the Tricomi initial guesses.
    const int n = 28;
    for (int i = 0; i < 14; ++i) {
        double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = pk;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        nodes[i] = x;
        weights[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

double es_norm()
{
    static const double value = [] {
        double nd[14], wt[14];
        gauss_legendre28_half(nd, wt);
        // Composite 28-point Gauss-Legendre over 512 panels of [-1, 1].
        const int panels = 512;
        double acc = 0.0;
        for (int p = 0; p < panels; ++p) {
            const double a = -1.0 + 2.0 * p / panels, b = -1.0 + 2.0 * (p + 1) / panels;
            const double mid = 0.5 * (a + b), half = 0.5 * (b - a);
            for (int i = 0; i < 14; ++i)
                acc += half * wt[i] * (es_kernel(mid + half * nd[i]) + es_kernel(mid - half * nd[i]));
        }
        return 4.0 * acc;
    }();
    return value;
}

std::vector<double> conv_corr(size_t image, size_t grid)
{
    double nd[14], wt[14];
    gauss_legendre28_half(nd, wt);
    const double norm = es_norm();
    std::vector<double> c(image / 2 + 1);
    for (size_t t = 0; t <= image / 2; ++t) {
        const double tn = static_cast<double>(t) / static_cast<double>(grid);
        double acc = 0.0;
        for (int i = 0; i < 14; ++i) acc += es_kernel(nd[i]) * wt[i] * std::cos(M_PI * tn * kSupport * nd[i]);
        c[t] = acc * kSupport / norm;
    }
    return c;
}

std::vector<float> corr_table(size_t image, size_t grid)
{
    const std::vector<double> c = conv_corr(image, grid);
    const double norm = es_norm();
    std::vector<float> out(c.size());
    for (size_t t = 0; t < c.size(); ++t) out[t] = static_cast<float>(1.0 / (c[t] * norm));
    return out;
}

}  // namespace fip

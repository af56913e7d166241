// Preprocessing on the GPU: SVD-plane uv projection and XX/YY collapse
// (SVD_MFS.py:38-110; SVDPreprocessKernels.cu:41-237).
//
// The projection reduces the baseline mean and covariance in FP64 with fixed
// per-block partial sums (summed on the host in block order, so the result is
// deterministic), diagonalises the 3x3 covariance on the host in FP64, and
// applies B = D.[V0 V1] as one coalesced streaming kernel.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_eigen.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace fip {
namespace {

constexpr double kLight = 299792458.0;   // SVD_MFS.py:7
constexpr int kPT = 256;                 // threads per block
constexpr int kPerThread = 16;           // samples per thread per block

// Sample n = row + chan*nrows (channel-major flattening, SVD_MFS.py:48-52).
__device__ __forceinline__ void sample(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                       size_t nrows, size_t n, double d[3])
{
    const size_t row = n % nrows, chan = n / nrows;
    const double sc = freqs[chan] / kLight;
    d[0] = uvw[3 * row] * sc;
    d[1] = uvw[3 * row + 1] * sc;
    d[2] = uvw[3 * row + 2] * sc;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh, double* out)
{
    for (int k = 0; k < NV; ++k) sh[k * kPT + threadIdx.x] = v[k];
    __syncthreads();
    for (int st = kPT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < NV; ++k) sh[k * kPT + threadIdx.x] += sh[k * kPT + threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) out[k] = sh[k * kPT];
}

// Per-block sums of D (3 values per block).
__global__ __launch_bounds__(kPT) void k_proj_sum(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                                  size_t nrows, size_t ns, double* __restrict__ part)
{
    __shared__ double sh[3 * kPT];
    double v[3] = {0, 0, 0};
    const size_t base = (size_t)blockIdx.x * kPT * kPerThread;
    for (int k = 0; k < kPerThread; ++k) {
        const size_t n = base + (size_t)k * kPT + threadIdx.x;
        if (n < ns) {
            double d[3];
            sample(uvw, freqs, nrows, n, d);
            v[0] += d[0];
            v[1] += d[1];
            v[2] += d[2];
        }
    }
    block_sum<3>(v, sh, part + 3 * blockIdx.x);
}

// Per-block centred second moments (xx, xy, xz, yy, yz, zz).
__global__ __launch_bounds__(kPT) void k_proj_cov(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                                  size_t nrows, size_t ns, double m0, double m1, double m2,
                                                  double* __restrict__ part)
{
    __shared__ double sh[6 * kPT];
    double v[6] = {0, 0, 0, 0, 0, 0};
    const size_t base = (size_t)blockIdx.x * kPT * kPerThread;
    for (int k = 0; k < kPerThread; ++k) {
        const size_t n = base + (size_t)k * kPT + threadIdx.x;
        if (n < ns) {
            double d[3];
            sample(uvw, freqs, nrows, n, d);
            const double a = d[0] - m0, b = d[1] - m1, c = d[2] - m2;
            v[0] += a * a;
            v[1] += a * b;
            v[2] += a * c;
            v[3] += b * b;
            v[4] += b * c;
            v[5] += c * c;
        }
    }
    block_sum<6>(v, sh, part + 6 * blockIdx.x);
}

// B[n] = (D_n - mean).Vj + mean.Vj (SVD_MFS.py:68, 105-110), streaming.
__global__ __launch_bounds__(kPT) void k_proj_apply(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                                    size_t nrows, size_t ns, double m0, double m1, double m2,
                                                    double v00, double v01, double v02, double v10, double v11,
                                                    double v12, float2* __restrict__ bin)
{
    const size_t n = (size_t)blockIdx.x * kPT + threadIdx.x;
    if (n >= ns) return;
    double d[3];
    sample(uvw, freqs, nrows, n, d);
    const double a = d[0] - m0, b = d[1] - m1, c = d[2] - m2;
    const double mv0 = m0 * v00 + m1 * v01 + m2 * v02, mv1 = m0 * v10 + m1 * v11 + m2 * v12;
    bin[n] = make_float2((float)(a * v00 + b * v01 + c * v02 + mv0), (float)(a * v10 + b * v11 + c * v12 + mv1));
}

// (w0 XX + w3 YY)/(w0 + w3), w = weight * !flag, 0 where the sum is not
// positive (SVDPreprocessKernels.cu:208-237).
__global__ __launch_bounds__(kPT) void k_collapse(const float* __restrict__ xr, const float* __restrict__ xi,
                                                  const float* __restrict__ yr, const float* __restrict__ yi,
                                                  const unsigned char* __restrict__ f0,
                                                  const unsigned char* __restrict__ f3,
                                                  const float* __restrict__ w0, const float* __restrict__ w3,
                                                  size_t n, float* __restrict__ ore, float* __restrict__ oim)
{
    const size_t i = (size_t)blockIdx.x * kPT + threadIdx.x;
    if (i >= n) return;
    const float a = f0[i] ? 0.0f : w0[i];
    const float b = f3[i] ? 0.0f : w3[i];
    const float s = a + b;
    if (s > 0.0f) {
        ore[i] = (xr[i] * a + yr[i] * b) / s;
        oim[i] = (xi[i] * a + yi[i] * b) / s;
    } else {
        ore[i] = 0.0f;
        oim[i] = 0.0f;
    }
}

}  // namespace
}  // namespace fip

using namespace fip;

extern "C" int fip_project(const double* uvw, size_t nrows, const double* freqs, size_t nchan, float* bin,
                           float* vin)
{
    try {
        if (!uvw || !freqs || !bin || !vin) return set_error(FIP_ERR_ARG, "null pointer argument");
        const size_t ns = nrows * nchan;
        if (nrows == 0 || nchan == 0 || ns < 2) return set_error(FIP_ERR_ARG, "need at least two samples");
        for (size_t c = 0; c < nchan; ++c)
            if (!std::isfinite(freqs[c])) return set_error(FIP_ERR_ARG, "non-finite frequency");
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        const size_t per_block = (size_t)kPT * kPerThread;
        const size_t nb = (ns + per_block - 1) / per_block;
        DevBuf duvw, dfreq, dpart, dbin;
        FIP_HIP(duvw.ensure(nrows * 3 * sizeof(double)));
        FIP_HIP(dfreq.ensure(nchan * sizeof(double)));
        FIP_HIP(dpart.ensure(nb * 6 * sizeof(double)));
        FIP_HIP(dbin.ensure(ns * sizeof(float2)));
        FIP_HIP(hipMemcpy(duvw.p, uvw, nrows * 3 * sizeof(double), hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(dfreq.p, freqs, nchan * sizeof(double), hipMemcpyHostToDevice));
        hipStream_t st = C->stream;
        std::vector<double> part(nb * 6);
        hipLaunchKernelGGL(k_proj_sum, dim3((unsigned)nb), dim3(kPT), 0, st, duvw.as<double>(), dfreq.as<double>(),
                           nrows, ns, dpart.as<double>());
        FIP_HIP(hipGetLastError());
        FIP_HIP(hipMemcpyAsync(part.data(), dpart.p, nb * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        FIP_HIP(hipStreamSynchronize(st));
        double mu[3] = {0, 0, 0};
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < 3; ++k) mu[k] += part[3 * b + k];
        for (double& m : mu) m /= (double)ns;
        hipLaunchKernelGGL(k_proj_cov, dim3((unsigned)nb), dim3(kPT), 0, st, duvw.as<double>(), dfreq.as<double>(),
                           nrows, ns, mu[0], mu[1], mu[2], dpart.as<double>());
        FIP_HIP(hipGetLastError());
        FIP_HIP(hipMemcpyAsync(part.data(), dpart.p, nb * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        FIP_HIP(hipStreamSynchronize(st));
        double s6[6] = {0, 0, 0, 0, 0, 0};
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < 6; ++k) s6[k] += part[6 * b + k];
        const double dn = (double)(ns - 1);
        const double cov[3][3] = {{s6[0] / dn, s6[1] / dn, s6[2] / dn},
                                  {s6[1] / dn, s6[3] / dn, s6[4] / dn},
                                  {s6[2] / dn, s6[4] / dn, s6[5] / dn}};
        double V[3][3];
        pca_basis(cov, V);
        hipLaunchKernelGGL(k_proj_apply, dim3((unsigned)((ns + kPT - 1) / kPT)), dim3(kPT), 0, st, duvw.as<double>(),
                           dfreq.as<double>(), nrows, ns, mu[0], mu[1], mu[2], V[0][0], V[0][1], V[0][2], V[1][0],
                           V[1][1], V[1][2], dbin.as<float2>());
        FIP_HIP(hipGetLastError());
        FIP_HIP(hipMemcpyAsync(bin, dbin.p, ns * sizeof(float2), hipMemcpyDeviceToHost, st));
        FIP_HIP(hipStreamSynchronize(st));
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) vin[3 * a + b] = (float)V[a][b];
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

extern "C" int fip_collapse(const float* xx_re, const float* xx_im, const float* yy_re, const float* yy_im,
                            const unsigned char* flag0, const unsigned char* flag3, const float* w0,
                            const float* w3, size_t n, float* out_re, float* out_im)
{
    try {
        if (!xx_re || !xx_im || !yy_re || !yy_im || !flag0 || !flag3 || !w0 || !w3 || !out_re || !out_im)
            return set_error(FIP_ERR_ARG, "null pointer argument");
        if (n == 0) return FIP_OK;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        DevBuf d;
        const size_t fb = n * sizeof(float);
        FIP_HIP(d.ensure(fb * 8 + 2 * n + 64));
        char* base = d.as<char>();
        float* xr = reinterpret_cast<float*>(base);
        float* xi = xr + n;
        float* yr = xi + n;
        float* yi = yr + n;
        float* a = yi + n;
        float* b = a + n;
        float* ore = b + n;
        float* oim = ore + n;
        unsigned char* f0 = reinterpret_cast<unsigned char*>(oim + n);
        unsigned char* f3 = f0 + n;
        FIP_HIP(hipMemcpy(xr, xx_re, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(xi, xx_im, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(yr, yy_re, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(yi, yy_im, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(a, w0, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(b, w3, fb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(f0, flag0, n, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(f3, flag3, n, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_collapse, dim3((unsigned)((n + kPT - 1) / kPT)), dim3(kPT), 0, C->stream, xr, xi, yr, yi,
                           f0, f3, a, b, n, ore, oim);
        FIP_HIP(hipGetLastError());
        FIP_HIP(hipStreamSynchronize(C->stream));
        FIP_HIP(hipMemcpy(out_re, ore, fb, hipMemcpyDeviceToHost));
        FIP_HIP(hipMemcpy(out_im, oim, fb, hipMemcpyDeviceToHost));
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

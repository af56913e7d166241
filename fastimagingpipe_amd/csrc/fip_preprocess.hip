// Preprocessing on the GPU: SVD-plane uv projection and XX/YY collapse
// (SVD_MFS.py:38-110; SVDPreprocessKernels.cu:41-237).
//
// The projection reduces the baseline mean and covariance in FP64 in one
// pass with fixed per-block partial sums (summed on the host in block order,
// so the result is deterministic), diagonalises the 3x3 covariance on the
// host in FP64 after one small read-back, and applies B = D.[V0 V1] as one
// coalesced streaming kernel.  fip_pipe_raw_device chains collapse,
// projection, imaging and trigger on the device.  Buffers are the context's
// grow-only ones: no call allocates once sizes are warm.
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
constexpr int kPerThread = 16;           // samples per thread per moment block
constexpr int kPerBlock = kPT * kPerThread;
constexpr int kMom = 9;                  // 3 first + 6 second moments
constexpr int kPar = 12;                 // per-snapshot projection parameters (mean, V0, V1, pad)

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

// One pass of moments for snapshot blockIdx.y: per block of kPerBlock
// samples, the sums of a = D - D_0 and of a a^T (xx, xy, xz, yy, yz, zz),
// with D_0 the snapshot's first sample (the shift keeps the one-pass
// covariance free of cancellation); block 0 also stores D_0 after the nb
// partials.  part: [S][nb + 1][kMom].
__global__ __launch_bounds__(kPT) void k_moments(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                                 size_t nrows, size_t ns, int nb, double* __restrict__ part)
{
    __shared__ double sh[kMom * kPT];
    const int s = blockIdx.y;
    const double* u = uvw + (size_t)s * nrows * 3;
    double d0[3];
    sample(u, freqs, nrows, 0, d0);
    double v[kMom] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const size_t base = (size_t)blockIdx.x * kPerBlock;
    for (int k = 0; k < kPerThread; ++k) {
        const size_t n = base + (size_t)k * kPT + threadIdx.x;
        if (n < ns) {
            double d[3];
            sample(u, freqs, nrows, n, d);
            const double a = d[0] - d0[0], b = d[1] - d0[1], c = d[2] - d0[2];
            v[0] += a;
            v[1] += b;
            v[2] += c;
            v[3] += a * a;
            v[4] += a * b;
            v[5] += a * c;
            v[6] += b * b;
            v[7] += b * c;
            v[8] += c * c;
        }
    }
    double* ps = part + (size_t)s * (nb + 1) * kMom;
    block_sum<kMom>(v, sh, ps + (size_t)blockIdx.x * kMom);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) ps[(size_t)nb * kMom + k] = d0[k];
}

// B[n] = (D_n - mean).Vj + mean.Vj (SVD_MFS.py:68, 105-110), streaming, for
// snapshot blockIdx.y into bin + s*stride.  par: [S][kPar] = mean, V0, V1.
__global__ __launch_bounds__(kPT) void k_proj_apply(const double* __restrict__ uvw, const double* __restrict__ freqs,
                                                    size_t nrows, size_t ns, const double* __restrict__ par,
                                                    float2* __restrict__ bin, size_t stride)
{
    const int s = blockIdx.y;
    const size_t n = (size_t)blockIdx.x * kPT + threadIdx.x;
    if (n >= ns) return;
    const double* q = par + (size_t)s * kPar;
    const double m0 = q[0], m1 = q[1], m2 = q[2];
    const double v00 = q[3], v01 = q[4], v02 = q[5], v10 = q[6], v11 = q[7], v12 = q[8];
    double d[3];
    sample(uvw + (size_t)s * nrows * 3, freqs, nrows, n, d);
    const double a = d[0] - m0, b = d[1] - m1, c = d[2] - m2;
    const double mv0 = m0 * v00 + m1 * v01 + m2 * v02, mv1 = m0 * v10 + m1 * v11 + m2 * v12;
    bin[(size_t)s * stride + n] =
        make_float2((float)(a * v00 + b * v01 + c * v02 + mv0), (float)(a * v10 + b * v11 + c * v12 + mv1));
}

// (w0 XX + w3 YY)/(w0 + w3), w = weight * !flag, 0 where the sum is not
// positive (SVDPreprocessKernels.cu:208-237), for snapshot blockIdx.y:
// inputs at s*n, outputs at s*ostride.
__global__ __launch_bounds__(kPT) void k_collapse(const float* __restrict__ xr, const float* __restrict__ xi,
                                                  const float* __restrict__ yr, const float* __restrict__ yi,
                                                  const unsigned char* __restrict__ f0,
                                                  const unsigned char* __restrict__ f3,
                                                  const float* __restrict__ w0, const float* __restrict__ w3,
                                                  size_t n, float* __restrict__ ore, float* __restrict__ oim,
                                                  size_t ostride)
{
    const size_t i = (size_t)blockIdx.x * kPT + threadIdx.x;
    if (i >= n) return;
    const size_t q = (size_t)blockIdx.y * n + i, o = (size_t)blockIdx.y * ostride + i;
    const float a = f0[q] ? 0.0f : w0[q];
    const float b = f3[q] ? 0.0f : w3[q];
    const float s = a + b;
    if (s > 0.0f) {
        ore[o] = (xr[q] * a + yr[q] * b) / s;
        oim[o] = (xi[q] * a + yi[q] * b) / s;
    } else {
        ore[o] = 0.0f;
        oim[o] = 0.0f;
    }
}

size_t moment_blocks(size_t ns) { return (ns + kPerBlock - 1) / kPerBlock; }

// Pinned layout of pp_host for S snapshots: partials read back, then the
// projection parameters and the float V matrices sent.
struct PpHostLayout {
    size_t part, par, vin, bytes;
};
PpHostLayout pp_host_layout(int S, size_t nb)
{
    PpHostLayout L;
    L.part = 0;
    L.par = L.part + (size_t)S * (nb + 1) * kMom * sizeof(double);
    L.vin = L.par + (size_t)S * kPar * sizeof(double);
    L.bytes = L.vin + (size_t)S * 9 * sizeof(float);
    return L;
}

// Projection of S snapshots whose uvw ([S][nrows][3]) and frequencies are in
// device memory: one moment pass, one read-back of the partials, the 3x3
// eigen-solve per snapshot on the host in FP64 (partials summed in block
// order, so the result is deterministic), then B into bin + s*stride
// (float2) and, when vin_dev is set, V as float into vin_dev + 9 s.
// Synchronises `st` once (after the moments); the rest stays enqueued.
int project_device(Context& C, const double* uvw, const double* freqs, size_t nrows, size_t nchan, int S,
                   float2* bin, size_t stride, float* vin_dev, double (*V)[3][3], hipStream_t st)
{
    const size_t ns = nrows * nchan;
    const size_t nb = moment_blocks(ns);
    const PpHostLayout L = pp_host_layout(S, nb);
    FIP_HIP(C.pp_part.ensure((size_t)S * (nb + 1) * kMom * sizeof(double)));
    FIP_HIP(C.pp_par.ensure((size_t)S * kPar * sizeof(double)));
    FIP_HIP(C.pp_host.ensure(L.bytes));
    char* h = static_cast<char*>(C.pp_host.p);
    hipLaunchKernelGGL(k_moments, dim3((unsigned)nb, (unsigned)S), dim3(kPT), 0, st, uvw, freqs, nrows, ns, (int)nb,
                       C.pp_part.as<double>());
    FIP_HIP(hipGetLastError());
    FIP_HIP(hipMemcpyAsync(h + L.part, C.pp_part.p, (size_t)S * (nb + 1) * kMom * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    FIP_HIP(hipStreamSynchronize(st));
    const double* part = reinterpret_cast<const double*>(h + L.part);
    double* par = reinterpret_cast<double*>(h + L.par);
    float* vf = reinterpret_cast<float*>(h + L.vin);
    for (int s = 0; s < S; ++s) {
        const double* ps = part + (size_t)s * (nb + 1) * kMom;
        double m[kMom] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t b = 0; b < nb; ++b)
            for (int k = 0; k < kMom; ++k) m[k] += ps[b * kMom + k];
        const double* d0 = ps + nb * kMom;
        const double dn = (double)ns;
        // mean = D_0 + sum(a)/n; centred sums = sum(a a^T) - n (mean-D_0)(mean-D_0)^T
        const double e[3] = {m[0] / dn, m[1] / dn, m[2] / dn};
        const double mu[3] = {d0[0] + e[0], d0[1] + e[1], d0[2] + e[2]};
        const double c00 = m[3] - dn * e[0] * e[0], c01 = m[4] - dn * e[0] * e[1], c02 = m[5] - dn * e[0] * e[2];
        const double c11 = m[6] - dn * e[1] * e[1], c12 = m[7] - dn * e[1] * e[2], c22 = m[8] - dn * e[2] * e[2];
        const double den = (double)(ns - 1);
        const double cov[3][3] = {{c00 / den, c01 / den, c02 / den},
                                  {c01 / den, c11 / den, c12 / den},
                                  {c02 / den, c12 / den, c22 / den}};
        pca_basis(cov, V[s]);
        double* q = par + (size_t)s * kPar;
        q[0] = mu[0];
        q[1] = mu[1];
        q[2] = mu[2];
        for (int k = 0; k < 3; ++k) {
            q[3 + k] = V[s][0][k];
            q[6 + k] = V[s][1][k];
        }
        q[9] = q[10] = q[11] = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) vf[9 * s + 3 * a + b] = (float)V[s][a][b];
    }
    FIP_HIP(hipMemcpyAsync(C.pp_par.p, h + L.par, (size_t)S * kPar * sizeof(double), hipMemcpyHostToDevice, st));
    if (vin_dev)
        FIP_HIP(hipMemcpyAsync(vin_dev, h + L.vin, (size_t)S * 9 * sizeof(float), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_proj_apply, dim3((unsigned)((ns + kPT - 1) / kPT), (unsigned)S), dim3(kPT), 0, st, uvw,
                       freqs, nrows, ns, C.pp_par.as<double>(), bin, stride);
    FIP_HIP(hipGetLastError());
    return FIP_OK;
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
        hipStream_t st = C->stream;
        // The pinned and device preprocessing buffers are shared with
        // fip_pipe_raw_device, whose last copies may still be in flight on a
        // caller's stream: wait for them first.
        if (int rc = C->order_work(st)) return rc;
        FIP_HIP(hipStreamSynchronize(st));
        FIP_HIP(C->pp_uvw.ensure(nrows * 3 * sizeof(double)));
        FIP_HIP(C->pp_freq.ensure(nchan * sizeof(double)));
        FIP_HIP(C->pp_in.ensure(ns * sizeof(float2)));
        FIP_HIP(hipMemcpy(C->pp_uvw.p, uvw, nrows * 3 * sizeof(double), hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->pp_freq.p, freqs, nchan * sizeof(double), hipMemcpyHostToDevice));
        double V[1][3][3];
        if (int rc = project_device(*C, C->pp_uvw.as<double>(), C->pp_freq.as<double>(), nrows, nchan, 1,
                                    C->pp_in.as<float2>(), ns, nullptr, V, st))
            return rc;
        FIP_HIP(hipMemcpyAsync(bin, C->pp_in.p, ns * sizeof(float2), hipMemcpyDeviceToHost, st));
        if (int rc = C->mark_work(st)) return rc;
        FIP_HIP(hipStreamSynchronize(st));
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) vin[3 * a + b] = (float)V[0][a][b];
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
        hipStream_t st = C->stream;
        if (int rc = C->order_work(st)) return rc;
        FIP_HIP(hipStreamSynchronize(st));
        const size_t fb = n * sizeof(float);
        FIP_HIP(C->pp_in.ensure(fb * 8 + 2 * n + 64));
        char* base = C->pp_in.as<char>();
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
        hipLaunchKernelGGL(k_collapse, dim3((unsigned)((n + kPT - 1) / kPT), 1), dim3(kPT), 0, st, xr, xi, yr, yi, f0,
                           f3, a, b, n, ore, oim, n);
        FIP_HIP(hipGetLastError());
        FIP_HIP(hipStreamSynchronize(st));
        FIP_HIP(hipMemcpy(out_re, ore, fb, hipMemcpyDeviceToHost));
        FIP_HIP(hipMemcpy(out_im, oim, fb, hipMemcpyDeviceToHost));
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

extern "C" int fip_pipe_raw_device(const fip_raw_view* raw, size_t image_size, float cell_size, size_t unit_size,
                                   float* result, float* bin_out, float* vin_out, void* stream)
{
    try {
        if (!raw || !result) return set_error(FIP_ERR_ARG, "null pointer argument");
        const fip_raw_view& r = *raw;
        if (!r.uvw || !r.freqs || !r.xx_re || !r.xx_im || !r.yy_re || !r.yy_im || !r.flag0 || !r.flag3 || !r.w0 ||
            !r.w3)
            return set_error(FIP_ERR_ARG, "null pointer in fip_raw_view");
        if (r.nsnap < 3) return set_error(FIP_ERR_NSNAP, "FIpipe needs at least 3 snapshots");
        if (r.nrows == 0 || r.nchan == 0) return set_error(FIP_ERR_ARG, "nrows and nchan must be positive");
        if (r.nrows > ((size_t)1 << 40) / r.nchan) return set_error(FIP_ERR_SIZE, "nrows * nchan too large");
        const size_t ns = r.nrows * r.nchan;
        if (ns < 2) return set_error(FIP_ERR_ARG, "need at least two samples per snapshot");
        if (int rc = check_shape(ns, image_size, unit_size, cell_size)) return rc;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : C->stream;
        const int S = 3;   // snapshots 0..2, as FIpipe (FIPkernels.cu:647)
        if (int rc = C->work.ensure(S, ns, P->t)) return rc;
        if (int rc = C->ensure_win012()) return rc;
        FIP_HIP(C->in_vr.ensure(S * ns * sizeof(float)));
        FIP_HIP(C->in_vi.ensure(S * ns * sizeof(float)));
        FIP_HIP(C->in_b.ensure(S * ns * sizeof(float2)));
        FIP_HIP(C->in_v.ensure(S * 9 * sizeof(float)));
        if (int rc = C->order_work(st)) return rc;
        hipLaunchKernelGGL(k_collapse, dim3((unsigned)((ns + kPT - 1) / kPT), (unsigned)S), dim3(kPT), 0, st, r.xx_re,
                           r.xx_im, r.yy_re, r.yy_im, r.flag0, r.flag3, r.w0, r.w3, ns, C->in_vr.as<float>(),
                           C->in_vi.as<float>(), ns);
        FIP_HIP(hipGetLastError());
        double V[S][3][3];
        if (int rc = project_device(*C, r.uvw, r.freqs, r.nrows, r.nchan, S, C->in_b.as<float2>(), ns,
                                    C->in_v.as<float>(), V, st))
            return rc;
        InputView in{C->in_vr.p, C->in_vi.p, C->in_b.as<float>(), C->in_v.as<float>(), nullptr, ns, (int)ns,
                     FIP_VIS_F32};
        if (int rc = pipe_sequence(C, P->t, C->work.view(ns), in, (int)ns, cell_size, C->tuning(image_size),
                                   C->win.as<int>(), (int)unit_size, result, st))
            return rc;
        if (bin_out)
            FIP_HIP(hipMemcpyAsync(bin_out, C->in_b.p, S * ns * sizeof(float2), hipMemcpyDeviceToDevice, st));
        if (vin_out) FIP_HIP(hipMemcpyAsync(vin_out, C->in_v.p, S * 9 * sizeof(float), hipMemcpyDeviceToDevice, st));
        if (int rc = C->mark_work(st)) return rc;
        if (C->profiling) {
            FIP_HIP(hipStreamSynchronize(st));
            return C->collect_stages();
        }
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

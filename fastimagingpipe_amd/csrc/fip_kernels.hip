// HIP kernels of the FIP-TOI imaging-and-trigger path for gfx950 (CDNA4).
//
// Per batch of S snapshots the path runs eight launches (each covers all S
// snapshots):
//   reset    zero histograms / per-snapshot scalars, FP64 re-projection geometry
//   bin      weighting (FIPkernels.cu:188-195), FP64 uv scaling, column-bucket
//            histogram with per-visibility slots
//   scan     exclusive scan of the bucket histogram
//   scatter  visibilities into bucket order (counting sort by grid column)
//   stage1   per uv column c >= 0: deterministic integer LDS gridding of the
//            Hermitian-folded grid column (FIPkernels.cu:197-245) and its
//            length-g inverse FFT along the row axis, cropped to I rows
//   stage2   per image row: complex-to-real length-g inverse FFT along the
//            column axis via a length-g/2 complex FFT, cropped, |.| and the
//            separable grid correction -> corrected image
//            (FIPkernels.cu:256-287), plus its maximum (spread slots)
//   splat    per target tile: affine + SIN re-projection (FP64 per group of
//            sources, FP32 within it) and the bilinear forward splat
//            accumulated exactly in int32 LDS, plus the image maximum
//            (FIPkernels.cu:289-497)
//   trigger  per tile and window of three images: tLISI (FIPkernels.cu:415-552)
//
// The image identity the stages rely on: with the reference's modulation
// (-1)^(r1+r2) in gridding and the (-1)^(|x|+|y|) flip after the FFT, the
// cropped real part is D[y][x] = Re sum_{r1,r2} Gc[r1][r2] e^{+2 pi i (r1 y + r2 x)/g}
// for the unmodulated grid Gc.  Folding Gc into the Hermitian grid
// Gh[r] = (Gc[r] + conj(Gc[-r]))/2 makes D = sum Gh e^{...} exactly, so only
// columns c = 0..g/2 of Gh are transformed and the second axis is a C2R.
#include "fip_fft.hpp"
#include "fip_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>

namespace fip {

// ---------------------------------------------------------------- helpers

// Spread-slot layout of the image maxima: [S][set 0..1][kMaxSlots] unsigned.
constexpr size_t kPmaxPer = (size_t)2 * kMaxSlots;   // unsigned per snapshot
__host__ __device__ inline size_t max_slot_index(int s, int set, int slot)
{
    return (size_t)s * kPmaxPer + (size_t)set * kMaxSlots + (size_t)slot;
}
__device__ __forceinline__ unsigned* max_slot(unsigned* pm, int s, int set, int slot)
{
    return pm + max_slot_index(s, set, slot);
}
size_t pmax_words(int S) { return (size_t)S * kPmaxPer; }
size_t pmax_index(int s, int set, int slot) { return max_slot_index(s, set, slot); }

// exp(beta*(sqrt(1-x^2)-1)) (FIPkernels.cu:67-71), 0 for |x| > 1, as
// 2^(C s - C) with C = beta log2 e and s = sqrt(1 - x^2) (1 - x^2 by one FMA,
// hardware sqrt and exp2, ~1 ulp each).  s <= 1 carries an absolute error of
// ~1e-7, so the exponent's error stays below ~2e-6 (C = 22.2) and the relative
// error of ES below ~1.5e-6 at every x.  The form without the subtraction,
// exp(-beta x^2 / (1 + s)), is no more accurate at this precision and costs
// a reciprocal and three more instructions per tap.
__device__ __forceinline__ float es_f(float x)
{
    constexpr float C = (float)(kBeta * 1.4426950408889634);
    const float q = fmaf(-x, x, 1.0f);
    if (q < 0.0f) return 0.0f;
    return __builtin_amdgcn_exp2f(fmaf(C, __builtin_amdgcn_sqrtf(q), -C));
}

template <class T> __device__ __forceinline__ float vis_to_f(const void* p, size_t i);
template <> __device__ __forceinline__ float vis_to_f<float>(const void* p, size_t i)
{
    return static_cast<const float*>(p)[i];
}
template <> __device__ __forceinline__ float vis_to_f<__hip_bfloat16>(const void* p, size_t i)
{
    return __bfloat162float(static_cast<const __hip_bfloat16*>(p)[i]);
}
template <> __device__ __forceinline__ float vis_to_f<__half>(const void* p, size_t i)
{
    return __half2float(static_cast<const __half*>(p)[i]);
}

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
// Max over the 64 lanes (all active), returned in every lane: DPP row shifts
// leave each row of 16 lanes' max in its last lane, and the four row maxima
// are combined as scalars (max is exact in any order).  10 instructions
// instead of 30 for the shuffle butterfly.
__device__ __forceinline__ unsigned wave_max_u(unsigned v)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));   // row_shr:1
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));   // row_shr:2
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));   // row_shr:4
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));   // row_shr:8
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 15);
    const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)v, 31);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 47);
    const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)v, 63);
    return max(max(a, b), max(c, d));
}
__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide max of a non-negative float (as ordered bits).  All threads call.
template <int NT>
__device__ __forceinline__ unsigned block_max_u(unsigned v, unsigned* sh)
{
    v = wave_max_u(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    unsigned r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) r = max(r, sh[i]);
    }
    return r;   // valid in thread 0
}

// Workgroups are dispatched round-robin over the 8 XCDs; this maps the
// hardware block id to a logical id so that each XCD works on one contiguous
// run of logical blocks (neighbouring rows/tiles then share an L2).
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    constexpr int nx = 8;
    const int per = nb / nx, rem = nb % nx;
    const int x = bid % nx, k = bid / nx;
    return x * per + min(x, rem) + k;
}

__device__ __forceinline__ long long to_fixed(float x)
{
    return (long long)rintf(x);   // |x| < 2^62 by construction; exact for |x| >= 2^24
}

// ----------------------------------------------------------------- reset

__global__ void k_reset(const float* __restrict__ vin, const int* __restrict__ nvis, int n_const,
                        const int* __restrict__ norm_n, int norm_const, SnapMeta* meta, SnapGeom* geom,
                        int* hist, unsigned* work, unsigned* pmax, int S, int G, int I, float cell)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S * G) hist[i] = 0;
    if (i < S * 2 * kMaxSlots) *max_slot(pmax, i / (2 * kMaxSlots), 0, i % (2 * kMaxSlots)) = 0u;
    if (i == 0) work[0] = 0u;
    if (i < S) {
        SnapMeta m;
        m.ngrid = 0;
        m.rmax = -1;
        m.r1max = -1;
        m.maxv_bits = 0u;
        m.maxp_bits = 0u;
        m.maxd_bits = 0u;
        m.nvis = nvis ? nvis[i] : n_const;
        m.norm_n = norm_n ? norm_n[i] : norm_const;
        meta[i] = m;

        // Re-projection geometry (coordschange FIPkernels.cu:289-299 and the
        // closed form of p2p :301-376), FP64 from the float V matrix.
        const float* V = vin + (size_t)i * 9;
        const double v00 = V[0], v01 = V[1], v10 = V[3], v11 = V[4];
        const double v20 = V[6], v21 = V[7], v22 = V[8];
        const double a22 = fabs(v22);
        SnapGeom gm;
        gm.m00 = -v00 / a22;
        gm.m01 = v10 / a22;
        gm.m10 = -v01 / a22;
        gm.m11 = v11 / a22;
        const double det = gm.m00 * gm.m11 - gm.m01 * gm.m10;
        gm.i00 = gm.m11 / det;
        gm.i01 = -gm.m01 / det;
        gm.i10 = -gm.m10 / det;
        gm.i11 = gm.m00 / det;
        gm.im00 = fabs(gm.m00) > 1e-12 ? 1.0 / gm.m00 : 0.0;
        gm.im10 = fabs(gm.m10) > 1e-12 ? 1.0 / gm.m10 : 0.0;
        const double xi = v20 / v22, eta = v21 / v22;
        const double delta = (double)cell;
        gm.sx = -xi / delta;
        gm.sy = eta / delta;
        gm.delta = delta;
        gm.sin_on = (xi * xi + eta * eta != 0.0) ? 1 : 0;
        // Singular values of the affine part, and the largest direction-cosine
        // radius any source or target can reach.
        const double h = (double)(I / 2);
        const double tr = gm.m00 * gm.m00 + gm.m01 * gm.m01 + gm.m10 * gm.m10 + gm.m11 * gm.m11;
        const double dd = fabs(det);
        const double disc = sqrt(fmax(0.0, tr * tr - 4.0 * dd * dd));
        const double smax = sqrt(0.5 * (tr + disc)), smin = sqrt(fmax(0.0, 0.5 * (tr - disc)));
        const double rmax = delta * (smax * h * 1.4142135623730951 + 4.0 + h);
        const double kap = sqrt(xi * xi + eta * eta) * rmax / sqrt(fmax(1e-300, 1.0 - rmax * rmax));
        gm.kappa = kap;
        if (!(rmax < 0.9) || !(kap < 0.5) || !isfinite(det) || det == 0.0) {
            gm.curv = -1.0;   // no usable bound: the splat scans whole images
        } else {
            gm.curv = kap;
        }
        // A pixel receives splat weight from the sources mapped into the open
        // 2 x 2 square around it.  The map expands by at least
        // smin (1 - kappa), so they lie in a set of diameter
        // 2 sqrt(2) / (smin (1 - kappa)), which holds at most (d + 1)^2
        // integer points.  Without a usable bound every pixel of the image
        // counts.
        {
            const double den = smin * (1.0 - kap);
            double cap = (gm.curv >= 0.0 && den > 0.0) ? 2.8284271247461903 / den + 1.0 : 0.0;
            cap = cap * cap;
            const double all = (double)I * (double)I;
            gm.ccap = (int)fmin(gm.curv >= 0.0 && den > 0.0 ? ceil(cap) : all, fmin(all, 1073741824.0));
            gm.ccap = max(gm.ccap, 1);
        }
        // n' = 1 - sqrt(1 - r^2) = r^2 (1/2 + r^2/8 + r^4/16 + 5 r^6/128 + ...);
        // for r^2 <= 0.01 the terms after r^8 are below 3e-12 of n'.
        // The splat takes this series along each group of 8 consecutive
        // sources of a row (k = 0..7) from its quadratic Taylor model in k.
        // r^2 is quadratic in k with |dr^2/dk| <= 2 r d and d2r^2/dk2 = 2 d^2,
        // d = cell |(m00, m10)|; for r <= 0.1 the series' third k-derivative
        // is at most 0.31 d^3, so the model's remainder at k <= 7 is at most
        // 17.6 d^3 in n', 17.6 d^3 max(|sx|, |sy|) in pixels.  The series
        // form is taken only where that is at most 5e-8 px (a tenth of an FP32
        // ulp of the splat's local positions); otherwise the closed form per
        // source.
        const double dcol = delta * sqrt(gm.m00 * gm.m00 + gm.m10 * gm.m10);
        const double qerr = 18.0 * dcol * dcol * dcol * fmax(fabs(gm.sx), fabs(gm.sy));
        gm.series = (rmax * rmax <= 0.01 && qerr <= 5e-8) ? 1 : 0;
        gm.pad = 0;
        geom[i] = gm;
    }
}

// ------------------------------------------------------------------- bin

// Threads of a bin workgroup and visibilities per thread (a workgroup bins
// kBinThreads * kBinVpt visibilities).  Large workgroups keep the number of
// per-workgroup bucket flushes (global atomics) low; 16 waves each keep the
// CUs busy although a step has only ~100 workgroups.
constexpr int kBinThreads = 1024;
constexpr int kBinVpt = 4;
// Largest grid size: bounds the bucket counts one bin or scan thread holds.
constexpr int kMaxGrid = 12288;

// Block-privatised column histogram: slots within the block come from LDS
// atomics, then one global atomic per non-empty bucket reserves the block's
// range, so the hot central buckets see one global atomic per block instead
// of one per visibility.
template <class VT>
__global__ __launch_bounds__(kBinThreads) void k_bin(InputView in, VisRec* __restrict__ rec, int2* __restrict__ bs,
                                             int* __restrict__ hist, SnapMeta* __restrict__ meta, int G,
                                             double scale)
{
    extern __shared__ int lh[];   // G ints: local counts, then block bases
    constexpr int NW = kBinThreads / 64;
    __shared__ int sh_i[3][NW];
    __shared__ unsigned sh_u[NW];
    const int s = blockIdx.y;
    const int ns = in.nvis ? in.nvis[s] : in.n_const;
    const int M = G / 2;
    for (int b = threadIdx.x; b < G; b += kBinThreads) lh[b] = 0;
    __syncthreads();
    int rmax = -1, r1max = -1, cnt = 0;
    unsigned vmax = 0u;
    int bk[kBinVpt], slot[kBinVpt];
    const float* V = in.vin + (size_t)s * 9;
    // computeVisWeighted: v / |V00 V11 - V01 V10| (FIPkernels.cu:188-195)
    const double J = fabs((double)V[0] * (double)V[4] - (double)V[1] * (double)V[3]);
    // Every input of the thread's visibilities is loaded before any is used,
    // so the loads are in flight together: unconditionally, from an index
    // clamped into this snapshot's rows (entries past its count are loaded
    // but not used), since a load behind a per-lane test would wait in turn.
    float lvr[kBinVpt], lvi[kBinVpt], lb1[kBinVpt], lb2[kBinVpt];
    const int nlast = max(ns - 1, 0);
#pragma unroll
    for (int k = 0; k < kBinVpt; ++k) {
        const int n = min(blockIdx.x * (kBinThreads * kBinVpt) + k * kBinThreads + threadIdx.x, nlast);
        const size_t q = (size_t)s * in.stride + (size_t)n;
        lvr[k] = vis_to_f<VT>(in.vis_re, q);
        lvi[k] = vis_to_f<VT>(in.vis_im, q);
        lb1[k] = in.bin[2 * q];
        lb2[k] = in.bin[2 * q + 1];
    }
    // Keeps the loads above the per-visibility tests (the compiler would
    // otherwise sink each into the branch that uses it).
#pragma unroll
    for (int k = 0; k < kBinVpt; ++k) asm volatile("" ::"v"(lvr[k]), "v"(lvi[k]), "v"(lb1[k]), "v"(lb2[k]));
#pragma unroll
    for (int k = 0; k < kBinVpt; ++k) {
        bk[k] = -1;
        slot[k] = 0;
        const int n = blockIdx.x * (kBinThreads * kBinVpt) + k * kBinThreads + threadIdx.x;
        if (n >= ns) continue;
        const size_t q = (size_t)s * in.stride + (size_t)n;
        const float vr = (float)((double)lvr[k] / J);
        const float vi = (float)((double)lvi[k] / J);
        // grid coordinates in FP64: B * (cell * g) (FIPkernels.cu:211-212, 584)
        const double p1 = (double)lb1[k] * scale;
        const double p2 = (double)lb2[k] * scale;
        // Coordinates far outside the grid have no taps (and must not overflow int).
        if (!(fabs(p1) < (double)G && fabs(p2) < (double)G)) continue;
        const double fb1 = floor(p1), fb2 = floor(p2);
        VisRec r;
        r.b1 = (int)fb1;
        r.f1 = (float)(p1 - fb1);
        r.b2 = (int)fb2;
        r.f2 = (float)(p2 - fb2);
        r.vr = vr;
        r.vi = vi;
        const int lo1 = max(r.b1 - 4 + (r.f1 > 0.0f ? 1 : 0), -M), hi1 = min(r.b1 + 4, M - 1);
        const int lo2 = max(r.b2 - 4 + (r.f2 > 0.0f ? 1 : 0), -M), hi2 = min(r.b2 + 4, M - 1);
        if (lo1 > hi1 || lo2 > hi2) continue;
        bk[k] = min(max(r.b2 + M, 0), G - 1);
        slot[k] = atomicAdd(&lh[bk[k]], 1);
        rec[q] = r;
        rmax = max(rmax, max(hi2, -lo2));
        r1max = max(r1max, max(hi1, -lo1));
        cnt += 1;
        vmax = max(vmax, __float_as_uint(fmaxf(fabsf(vr), fabsf(vi))));
    }
    __syncthreads();
    // Flush: each thread's bucket counts are read, then all of its global
    // atomics are issued before the first result is used.
    {
        constexpr int FL = (kMaxGrid + kBinThreads - 1) / kBinThreads;
        int fc[FL];
#pragma unroll
        for (int k = 0; k < FL; ++k) {
            const int b = threadIdx.x + k * kBinThreads;
            fc[k] = b < G ? lh[b] : 0;
        }
#pragma unroll
        for (int k = 0; k < FL; ++k) {
            const int b = threadIdx.x + k * kBinThreads;
            if (fc[k] > 0) fc[k] = atomicAdd(&hist[(size_t)s * G + b], fc[k]);
        }
#pragma unroll
        for (int k = 0; k < FL; ++k) {
            const int b = threadIdx.x + k * kBinThreads;
            if (b < G) lh[b] = fc[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBinVpt; ++k) {
        const int n = blockIdx.x * (kBinThreads * kBinVpt) + k * kBinThreads + threadIdx.x;
        if (n >= ns) continue;
        const size_t q = (size_t)s * in.stride + (size_t)n;
        bs[q] = bk[k] >= 0 ? make_int2(bk[k], lh[bk[k]] + slot[k]) : make_int2(-1, 0);
    }
    rmax = wave_max_i(rmax);
    r1max = wave_max_i(r1max);
    cnt = wave_sum_i(cnt);
    vmax = wave_max_u(vmax);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sh_i[0][wid] = rmax;
        sh_i[1][wid] = r1max;
        sh_i[2][wid] = cnt;
        sh_u[wid] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < NW; ++k) {
            rmax = max(rmax, sh_i[0][k]);
            r1max = max(r1max, sh_i[1][k]);
            vmax = max(vmax, sh_u[k]);
            cnt += sh_i[2][k];
        }
        if (cnt > 0) {
            atomicMax(&meta[s].rmax, rmax);
            atomicMax(&meta[s].r1max, r1max);
            atomicMax(&meta[s].maxv_bits, vmax);
            atomicAdd(&meta[s].ngrid, cnt);
        }
    }
}

// ------------------------------------------------------------------ scan

__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ hist, int* __restrict__ bstart, int G)
{
    // Exclusive prefix of the bucket counts: each thread loads a contiguous
    // chunk into registers (one round of loads), waves scan their chunk sums
    // with shuffles, one wave scans the 16 wave totals.
    constexpr int PER = (kMaxGrid + 1023) / 1024;
    __shared__ int wsum[16];
    const int s = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int per = (G + 1023) / 1024;
    const int beg = min(t * per, G), end = min(beg + per, G);
    const int* h = hist + (size_t)s * G;
    int* o = bstart + (size_t)s * (G + 1);
    int v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) v[k] = (beg + k < end) ? h[beg + k] : 0;
    int local = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) local += v[k];
    int x = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(w, d);
            if (lane >= d) w += y;
        }
        if (lane < 16) wsum[lane] = w;   // inclusive totals
    }
    __syncthreads();
    int run = x - local + (wid > 0 ? wsum[wid - 1] : 0);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (beg + k < end) o[beg + k] = run;
        run += v[k];
    }
    if (t == 1023) o[G] = wsum[15];
}

// --------------------------------------------------------------- scatter

__global__ __launch_bounds__(256) void k_scatter(const VisRec* __restrict__ rec, const int2* __restrict__ bs,
                                                 const int* __restrict__ bstart, VisRec* __restrict__ sorted,
                                                 const SnapMeta* __restrict__ meta, size_t stride, int G)
{
    const int s = blockIdx.y;
    const int n = blockIdx.x * 256 + threadIdx.x;
    // The last workgroup reaches past the largest count; n < stride keeps
    // the loads inside this snapshot's rows, so they are issued before its
    // count is known and all three are in flight together.
    if ((size_t)n >= stride) return;
    const size_t q = (size_t)s * stride + (size_t)n;
    const int2 key = bs[q];
    const VisRec r = rec[q];
    if (n >= meta[s].nvis || key.x < 0) return;
    sorted[(size_t)s * stride + (size_t)(bstart[(size_t)s * (G + 1) + key.x] + key.y)] = r;
}

// ---------------------------------------------------------------- stage 1

// One visibility record prepared for a column: the column weight
// ES((col - p2)/4)/2 (the Hermitian fold's 1/2) and the scale are folded
// into its value; the row taps are applied per LDS band.
struct ColRec {
    int b1;
    float f1, vr, vi;
    bool mirror, live;
};

// Taps follow FIPkernels.cu:213-243: r in [ceil(p-4), floor(p+4)] clipped to
// [-g/2, g/2-1].  A mirrored record contributes conj(v) at row -r.
__device__ __forceinline__ ColRec prep_rec(const VisRec& r, int col, bool mirror, int M, float fscale)
{
    ColRec c;
    c.b1 = r.b1;
    c.f1 = r.f1;
    c.mirror = mirror;
    const int lo2 = max(r.b2 - 4 + (r.f2 > 0.0f ? 1 : 0), -M);
    const int hi2 = min(r.b2 + 4, M - 1);
    c.live = (col >= lo2 && col <= hi2);
    const float w2 = c.live ? es_f(((float)(col - r.b2) - r.f2) * 0.25f) * 0.5f * fscale : 0.0f;
    c.vr = r.vr * w2;
    c.vi = (mirror ? -r.vi : r.vi) * w2;
    return c;
}

// Grid rows are grouped in at most kScaleBlocks blocks of at least 16 rows
// (by the index r mod G of row r); each block has its own power-of-two scale.
constexpr int kScaleBlocks = 96;

__device__ __forceinline__ int tap_row(const ColRec& c, int t1, int M)
{
    return c.mirror ? ((t1 == -M) ? -M : -t1) : t1;
}

// floor(x) and floor(x + 0.5) as int32 in one instruction each (the compiler
// keeps floor and the conversion separate without finite-math flags).
__device__ __forceinline__ int cvt_floor_i32(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ int cvt_round_half_up_i32(float x)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// ES inside its support, |x| <= 1 (as es_f without the range test).
__device__ __forceinline__ float es_in(float x)
{
    constexpr float C = (float)(kBeta * 1.4426950408889634);
    return __builtin_amdgcn_exp2f(fmaf(C, __builtin_amdgcn_sqrtf(fmaf(-x, x, 1.0f)), -C));
}

// Row taps of one prepared record, accumulated as integers into the LDS
// column (re, im interleaved; row r at padded index of r mod G), each scaled
// by its block's power of two.  A mirrored record adds at row -t1 (mod G,
// which also maps the edge tap t1 = -g/2 onto itself).  The tap offsets
// (t1 - p1)/4 step by 1/4 from the first tap; each contribution is rounded
// half up to an integer.  rbs is a power of two.
__device__ __forceinline__ void grid_taps(const ColRec& c, int M, int G, int rbs, const float* sc, int* col)
{
    const int lo1 = max(c.b1 - 4 + (c.f1 > 0.0f ? 1 : 0), -M);
    const int hi1 = min(c.b1 + 4, M - 1);
    float x = ((float)(lo1 - c.b1) - c.f1) * 0.25f;
    const int sgn = c.mirror ? -1 : 1;
    int row = sgn * lo1;   // the tap row sgn * t1, stepped with t1
    // The taps' rows are one contiguous run of r mod G, so they reach at most
    // two scale blocks: the first tap's and the last tap's.  Both scales are
    // read once; each tap selects its own, so no tap waits on an LDS read
    // (which would also wait for the previous taps' atomics).
    const int ra = row, rz = sgn * hi1;
    const unsigned ba = (unsigned)(ra < 0 ? ra + G : ra) / (unsigned)rbs;
    const unsigned bz = (unsigned)(rz < 0 ? rz + G : rz) / (unsigned)rbs;
    const float sca = sc[ba], scz = sc[bz];
    auto tap = [&]() {
        const int idx = row < 0 ? row + G : row;
        const int q = 2 * lds_pad(idx);
        const float wk = es_in(x) * ((unsigned)idx / (unsigned)rbs == ba ? sca : scz);
        atomicAdd(&col[q], cvt_round_half_up_i32(c.vr * wk));
        atomicAdd(&col[q + 1], cvt_round_half_up_i32(c.vi * wk));
        x += 0.25f;
        row += sgn;
    };
    // 8 taps (9 where the position is an integer), fewer only at the grid's
    // edge: the first eight need no bound test (3% faster at 4096^2, 7% at
    // 2048^2 with the ES form above, than one test per tap).
    const int n = hi1 - lo1 + 1;
    if (n >= 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) tap();
        if (n == 9) tap();
    } else {
        for (int j = 0; j < n; ++j) tap();
    }
}

// Adds the record's magnitude bound (units of amax / 4096, rounded up) to the
// mass of every scale block its taps touch (at most two).  The mass is 64-bit:
// a column holds up to two records per visibility (its direct and mirrored
// buckets), so a 32-bit sum of up to 4096 units per record would wrap after
// about 2^19 visibilities at one uv point.
__device__ __forceinline__ void tap_mass(const ColRec& c, int M, int G, int rbs, float qs, unsigned long long* mass)
{
    const int lo1 = max(c.b1 - 4 + (c.f1 > 0.0f ? 1 : 0), -M);
    const int hi1 = min(c.b1 + 4, M - 1);
    if (lo1 > hi1) return;
    const unsigned long long q = (unsigned long long)ceilf(fmaxf(fabsf(c.vr), fabsf(c.vi)) * qs);
    const int ra = tap_row(c, lo1, M), rb = tap_row(c, hi1, M);
    const int ba = (ra < 0 ? ra + G : ra) / rbs, bb = (rb < 0 ? rb + G : rb) / rbs;
    atomicAdd(&mass[ba], q);
    if (bb != ba) atomicAdd(&mass[bb], q);
}

// Record ranges of column c: direct buckets cc-4..cc+4 and mirrored buckets
// cm-4..cm+4 (x, y: direct [x, y); z, w: mirrored [z, w)).
template <int G>
__device__ __forceinline__ int4 column_ranges(const int* __restrict__ bst, int c)
{
    constexpr int M = G / 2;
    const int cc = (c == M) ? -M : c;
    const int cm = (c == 0) ? 0 : ((c == M) ? -M : -c);
    int4 r;
    r.x = bst[min(max(cc - 4 + M, 0), G - 1)];
    r.y = bst[min(max(cc + 4 + M, 0), G - 1) + 1];
    r.z = bst[min(max(cm - 4 + M, 0), G - 1)];
    r.w = bst[min(max(cm + 4 + M, 0), G - 1) + 1];
    return r;
}

// Threads of a stage-1 workgroup: the column FFT's own count, or 512 where
// that count is below 256 (g <= 1536), so that a column with many records
// still spreads its gridding over eight waves; the extra threads idle
// through the FFT stages.  (A floor of 512 for g = 3072 measured slower.)
template <int G>
constexpr int stage1_threads()
{
    return FftPlan<G>::nt < 256 ? 512 : FftPlan<G>::nt;
}

template <int G, int KR>
__device__ __forceinline__ void stage1_column(int s, int c, int4 rg, const VisRec* __restrict__ sorted,
                                              const SnapMeta& m,
                                              const float2* __restrict__ tw,
                                              float2* __restrict__ kt, size_t stride, int I, float2* fb,
                                              unsigned long long* mass, float* sc, float* isc)
{
    constexpr int NT = stage1_threads<G>();
    constexpr int M = G / 2;
    const int t = threadIdx.x;

    const VisRec* vs = sorted + (size_t)s * stride;
    const int cc = (c == M) ? -M : c;                        // direct column (centred)
    const int cm = (c == 0) ? 0 : ((c == M) ? -M : -c);      // mirrored column
    const int d0 = rg.x, d1 = rg.y, m0 = rg.z, m1 = rg.w;
    const int nd = d1 - d0, ntot = nd + (m1 - m0);

    // Exact integer accumulation with one power-of-two scale per block of
    // rows.  Pass 1 adds, per block, an upper bound of the magnitudes of the
    // records whose taps reach it (integers, so the bound does not depend on
    // the order of the LDS atomics); the scale keeps every int32 cell sum of
    // the block below 2^29 in magnitude.  Pass 2 accumulates the scaled,
    // rounded contributions; sums of integers are exact, so the column does
    // not depend on the order of the atomics either.
    const float maxv = __uint_as_float(m.maxv_bits);
    const float amax = fmaxf(0.5f * maxv, 1e-30f);   // |v| ES / 2 <= amax
    const float qs = 4096.0f / amax;
    // At least 16 rows per block, so the 9 taps of a record reach at most two
    // blocks (two ends of a contiguous run of r mod G); at most kScaleBlocks.
    constexpr int RBS = (G / kScaleBlocks > 16) ? G / kScaleBlocks : 16;
    static_assert(G % RBS == 0 && G / RBS <= kScaleBlocks, "scale blocks");

    // Direct records (buckets cc-4..cc+4) then mirrored ones (cm-4..cm+4) in
    // one numbering; each thread holds KR of them in registers at a time.
    ColRec cr[KR];
    auto load = [&](int base) {
        VisRec r[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int n = base + k * NT + t;
            if (n < ntot) r[k] = vs[n < nd ? d0 + n : m0 + (n - nd)];
        }
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int n = base + k * NT + t;
            if (n < ntot)
                cr[k] = prep_rec(r[k], n < nd ? cc : cm, n >= nd, M, 1.0f);
            else
                cr[k].live = false;
        }
    };
    const bool resident = ntot <= KR * NT;
    if (resident) load(0);
    int* col = reinterpret_cast<int*>(fb);
    for (int i = t; i < G; i += NT) fb[lds_pad(i)] = make_float2(0.0f, 0.0f);   // integer zero too
    for (int i = t; i < kScaleBlocks; i += NT) mass[i] = 0ull;
    __syncthreads();
    for (int base = 0; base < ntot; base += KR * NT) {
        if (!resident) load(base);
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (cr[k].live) tap_mass(cr[k], M, G, RBS, qs, mass);
    }
    __syncthreads();
    for (int i = t; i < kScaleBlocks; i += NT) {
        // bound = mass * amax / 4096 (+ one unit per record of rounding slack
        // is already inside the ceil); 2^e > bound.
        const float bound = (float)mass[i] * (amax / 4096.0f) * 1.0001f;
        int e = 0;
        frexpf(fmaxf(bound, 1e-30f), &e);
        const int k = min(29 - e, 120);   // empty blocks: any finite scale
        sc[i] = ldexpf(1.0f, k);
        isc[i] = ldexpf(1.0f, -k);
    }
    __syncthreads();
    for (int base = 0; base < ntot; base += KR * NT) {
        if (!resident) load(base);
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (cr[k].live) grid_taps(cr[k], M, G, RBS, sc, col);
    }
    __syncthreads();
    const int rlo = -m.r1max, rhi = min(m.r1max, M - 1);
    for (int i = rlo + t; i <= rhi; i += NT) {
        const int idx = i < 0 ? i + G : i;
        const int q = lds_pad(idx);
        const int re = col[2 * q], im = col[2 * q + 1];
        const float k = isc[idx / RBS];
        fb[q] = make_float2((float)re * k, (float)im * k);
    }
    __syncthreads();
    fft_inverse<G>(fb, tw, t);
    // Crop y in [-h, h).  The row correction 1/(c(|y|) norm) is a factor per
    // image row, applied by the row pass (one load per row there instead of
    // one per row and column here).
    const int h = I / 2;
    float2* ks = kt + (size_t)s * ((size_t)I * (M + 1));
    // Two consecutive rows per thread and one 16-byte store: with yp even the
    // pair shares a K-tile run (kKTile is even) at a 16-byte offset, and its
    // LDS slots are adjacent (an even index and the next one never straddle
    // a padding slot; h and G are even, so y and its wrap keep the parity).
    static_assert(kKTile % 2 == 0, "row pairs inside a K-tile run");
#pragma unroll 2
    for (int yp = 2 * t; yp < I; yp += 2 * NT) {
        const int y = yp - h;
        const int q = lds_pad(y < 0 ? y + G : y);
        const float2 v0 = fb[q], v1 = fb[q + 1];
        *reinterpret_cast<float4*>(ks + ((size_t)(yp / kKTile) * (M + 1) + c) * kKTile + (yp % kKTile)) =
            make_float4(v0.x, v0.y, v1.x, v1.y);
    }
}

template <int G, int KR>
__global__ __launch_bounds__(stage1_threads<G>(), (stage1_threads<G>() >= 1024 ? 1 : 2)) void k_stage1(const VisRec* __restrict__ sorted,
                                                            const int* __restrict__ bstart,
                                                            const SnapMeta* __restrict__ meta,
                                                            const float2* __restrict__ tw,
                                                            float2* __restrict__ kt, size_t stride, int I)
{
    __shared__ float2 fb[lds_len(G)];
    __shared__ unsigned long long mass[kScaleBlocks];
    __shared__ float sc[kScaleBlocks], isc[kScaleBlocks];
    // Column-major over snapshots: the heavily populated low columns of
    // every snapshot are dispatched first, so none of them ends the launch.
    const int S = (int)(gridDim.x / (G / 2 + 1));
    const int s = (int)(blockIdx.x % S);
    const int c = (int)(blockIdx.x / S);
    // The bucket ranges do not depend on the metadata: both loads are in
    // flight together.
    const int4 rg = column_ranges<G>(bstart + (size_t)s * (G + 1), c);
    const SnapMeta m = meta[s];
    if (c > m.rmax) return;
    stage1_column<G, KR>(s, c, rg, sorted, m, tw, kt, stride, I, fb, mass, sc, isc);
}

// Persistent form: a fixed grid pulls (snapshot, column) items from a
// counter until all S*(M+1) items are claimed; every workgroup exits.
template <int G, int KR>
__global__ __launch_bounds__(stage1_threads<G>()) void k_stage1_persistent(
    const VisRec* __restrict__ sorted, const int* __restrict__ bstart, const SnapMeta* __restrict__ meta,
    const float2* __restrict__ tw, float2* __restrict__ kt, size_t stride,
    int I, int S, unsigned* work)
{
    __shared__ float2 fb[lds_len(G)];
    __shared__ unsigned long long mass[kScaleBlocks];
    __shared__ float sc[kScaleBlocks], isc[kScaleBlocks];
    __shared__ unsigned item;
    constexpr unsigned per = G / 2 + 1;
    const unsigned total = (unsigned)S * per;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(work, 1u);
        __syncthreads();
        const unsigned it = item;
        __syncthreads();
        if (it >= total) break;
        const int s = (int)(it % (unsigned)S), c = (int)(it / (unsigned)S);
        const int4 rg = column_ranges<G>(bstart + (size_t)s * (G + 1), c);
        const SnapMeta m = meta[s];
        if (c > m.rmax) continue;
        stage1_column<G, KR>(s, c, rg, sorted, m, tw, kt, stride, I, fb, mass, sc, isc);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- stage 2

// Gather and C2R split of ROWS image rows in one pass: Z[k] = E[k] + i O[k],
// E = K[k] + K~[k+M], O = (K[k] - K~[k+M]) e^{2 pi i k/G},
// K~[k+M] = conj(K[M-k]) (k > 0), K[M] (k = 0).  A thread takes the pairs
// (p, M-p) of every row, reads K[p] and K[M-p] straight from the K tiles
// (zero beyond cmax) and writes Z[p], Z[M-p]; the pairs cover every index
// 0..M-1.  wpost[M - p] = -conj(wpost[p]).  The loads of all the thread's
// pairs are issued together (load), then paired into LDS (store).
template <int M, int ROWS>
struct C2RGather {
    static constexpr int NT = FftPlan<M>::nt * ROWS;          // threads of the workgroup
    static constexpr int PB = (M / 2 + NT) / NT;               // pairs per thread: ceil((M/2+1)/NT)
    float2 a[PB][ROWS], b[PB][ROWS], wv[PB];

    // kt0: K tile run of the first row of the group (row y0 of its snapshot).
    __device__ __forceinline__ void load(const float2* __restrict__ kt0, const float2* __restrict__ wpost,
                                         int cmax, int tid)
    {
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int p = tid + u * NT;
            const bool ok = p <= M / 2;
            wv[u] = ok ? wpost[p] : make_float2(0.0f, 0.0f);
            if constexpr (ROWS % 2 == 0) {
                // Two rows per 16-byte load: y0 is a multiple of ROWS, so the
                // pair is 16-byte aligned inside its 64-byte tile run.
#pragma unroll
                for (int r = 0; r < ROWS; r += 2) {
                    const float4 va = (ok && p <= cmax)
                                          ? *reinterpret_cast<const float4*>(kt0 + (size_t)p * kKTile + r)
                                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const float4 vb = (ok && M - p <= cmax)
                                          ? *reinterpret_cast<const float4*>(kt0 + (size_t)(M - p) * kKTile + r)
                                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    a[u][r] = make_float2(va.x, va.y);
                    a[u][r + 1] = make_float2(va.z, va.w);
                    b[u][r] = make_float2(vb.x, vb.y);
                    b[u][r + 1] = make_float2(vb.z, vb.w);
                }
            } else {
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    a[u][r] = (ok && p <= cmax) ? kt0[(size_t)p * kKTile + r] : make_float2(0.0f, 0.0f);
                    b[u][r] = (ok && M - p <= cmax) ? kt0[(size_t)(M - p) * kKTile + r] : make_float2(0.0f, 0.0f);
                }
            }
        }
    }

    __device__ __forceinline__ void store(float2* zb, int tid) const
    {
        constexpr int L = lds_len(M + 1);
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int p = tid + u * NT;
            if (p > M / 2) break;
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                float2* zr = zb + r * L;
                const float2 av = a[u][r], bv = b[u][r];
                if (p == 0) {
                    zr[lds_pad(0)] = cadd(cadd(av, bv), cmul_i(csub(av, bv)));
                } else {
                    const float2 kp = cconj(bv);   // K~[p+M]
                    zr[lds_pad(p)] = cadd(cadd(av, kp), cmul_i(cmul(csub(av, kp), wv[u])));
                    if (p != M - p) {
                        const float2 km = cconj(av);   // K~[M-p+M]
                        const float2 wm = make_float2(-wv[u].x, wv[u].y);
                        zr[lds_pad(M - p)] = cadd(cadd(bv, km), cmul_i(cmul(csub(bv, km), wm)));
                    }
                }
            }
        }
    }
};

// The same gather for a K band narrower than half the row (cmax < M/2), the
// usual case (the uv coverage spans under half the Nyquist band): every
// K[M-p] of a pair is then zero, so only K[p] is read and the pairing keeps
// the operations on the non-zero terms, Z[p] = K[p] + i K[p] w_p and
// Z[M-p] = conj(K[p]) + i (-conj(K[p])) w_{M-p}.  Pair blocks that start
// beyond cmax only write zeros.
template <int M, int ROWS>
struct C2RGatherLow {
    static constexpr int NT = FftPlan<M>::nt * ROWS;
    static constexpr int PB = (M / 2 + NT) / NT;
    float2 a[PB][ROWS], wv[PB];

    __device__ __forceinline__ void load(const float2* __restrict__ kt0, const float2* __restrict__ wpost,
                                         int cmax, int tid)
    {
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            if (u * NT > cmax) break;
            const int p = tid + u * NT;
            const bool ok = p <= cmax;
            wv[u] = ok ? wpost[p] : make_float2(0.0f, 0.0f);
            if constexpr (ROWS % 2 == 0) {
#pragma unroll
                for (int r = 0; r < ROWS; r += 2) {
                    const float4 va = ok ? *reinterpret_cast<const float4*>(kt0 + (size_t)p * kKTile + r)
                                         : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    a[u][r] = make_float2(va.x, va.y);
                    a[u][r + 1] = make_float2(va.z, va.w);
                }
            } else {
#pragma unroll
                for (int r = 0; r < ROWS; ++r)
                    a[u][r] = ok ? kt0[(size_t)p * kKTile + r] : make_float2(0.0f, 0.0f);
            }
        }
    }

    __device__ __forceinline__ void store(float2* zb, int cmax, int tid) const
    {
        constexpr int L = lds_len(M + 1);
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int p = tid + u * NT;
            if (p > M / 2) break;
            if (u * NT > cmax) {
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    float2* zr = zb + r * L;
                    zr[lds_pad(p)] = make_float2(0.0f, 0.0f);
                    zr[lds_pad(M - p)] = make_float2(0.0f, 0.0f);
                }
                continue;
            }
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
                float2* zr = zb + r * L;
                const float2 av = a[u][r];
                if (p == 0) {
                    zr[lds_pad(0)] = cadd(av, cmul_i(av));
                } else {
                    zr[lds_pad(p)] = cadd(av, cmul_i(cmul(av, wv[u])));
                    if (p != M - p) {
                        const float2 km = cconj(av);
                        const float2 nk = make_float2(-av.x, av.y);   // -conj(K[p])
                        const float2 wm = make_float2(-wv[u].x, wv[u].y);
                        zr[lds_pad(M - p)] = cadd(km, cmul_i(cmul(nk, wm)));
                    }
                }
            }
        }
    }
};

// First K tile run of image row y0 of snapshot s.
template <int M>
__device__ __forceinline__ const float2* c2r_rows(const float2* __restrict__ kt, int s, int y0, int I)
{
    const float2* ks = kt + (size_t)s * ((size_t)I * (M + 1));
    return ks + (size_t)(y0 / kKTile) * (M + 1) * kKTile + (y0 % kKTile);
}

// d[2n] = Re z[n], d[2n+1] = Im z[n]; crop x in [-h, h), |.| and the grid
// correction 1/(c(|x|) c(|y|) norm^2) (FIPkernels.cu:256-287) for all ROWS
// rows of a workgroup at once:
// each thread takes whole columns, so a column's correction factor is loaded
// once for all the rows; returns the thread's max.  out0: output row of the
// workgroup's first row.
template <int M, int ROWS, int NTB>
__device__ __forceinline__ float c2r_output_rows(const float2* zb, const float* __restrict__ corr, float* out0,
                                                 int y0, int I, int tid)
{
    constexpr int G = 2 * M;
    constexpr int L = lds_len(M + 1);
    const int h = I / 2;
    float lmax = 0.0f;
    // Row correction of each row (the grid correction is separable).
    float cy[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = y0 + r - h;
        cy[r] = corr[y < 0 ? -y : y];
    }
    for (int c0 = tid; c0 < I; c0 += 4 * NTB) {
        float cv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int x = c0 + u * NTB - h;
            cv[u] = (c0 + u * NTB < I) ? corr[x < 0 ? -x : x] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int col = c0 + u * NTB;
            if (col < I) {
                const int x = col - h;
                const int mm = x < 0 ? x + G : x;
#pragma unroll
                for (int r = 0; r < ROWS; ++r) {
                    const float2 v = zb[r * L + lds_pad(mm >> 1)];
                    const float val = fabsf((mm & 1) ? v.y : v.x) * (cv[u] * cy[r]);
                    // Non-temporal store: 122-127 -> 108-109 us at 4096^2, and
                    // the splat that reads these images next is not slower.
                    __builtin_nontemporal_store(val, out0 + (size_t)r * I + col);
                    lmax = fmaxf(lmax, val);
                }
            }
        }
    }
    return lmax;
}

// The image size of a plan: G = 2M = 1.5 I.
template <int M>
constexpr int c2r_image() { return 4 * M / 3; }

// Whether c2r_output_fixed applies: each of a thread's columns lies wholly on
// one side of the crop's wrap point h, and every column's offset from the
// thread's first one is a multiple of 64.
template <int M, int NTB>
constexpr bool c2r_fixed_ok()
{
    constexpr int G = 2 * M, I = c2r_image<M>(), h = I / 2;
    return 3 * I == 2 * G && NTB % 64 == 0 && h % 64 == 0 && (G - h) % 64 == 0 && NTB <= h && h % NTB == 0;
}

// c2r_output_rows for the plan's own image size, with the column walk resolved
// at compile time: thread t takes columns t + j*NTB.  Column col maps to the
// real sample mm = col + G - h (col < h) or col - h (col >= h), i.e. t plus a
// constant C_j that is a multiple of 64, so its float in the LDS row,
// 2*lds_pad(mm >> 1) + (mm & 1), is t + 2*(t >> 6) + C_j + C_j/32: one 4-byte
// LDS read at an immediate offset from a per-thread base, in place of an
// index computation, an 8-byte read and a select.  The same values, products
// and maxima as c2r_output_rows.
template <int M, int ROWS, int NTB>
__device__ __forceinline__ float c2r_output_fixed(const float2* zb, const float* __restrict__ corr, float* out0,
                                                  int y0, int tid)
{
    constexpr int G = 2 * M, I = c2r_image<M>(), h = I / 2;
    constexpr int L2 = 2 * lds_len(M + 1);   // floats per LDS row
    constexpr int NC = I / NTB;              // columns per thread
    float lmax = 0.0f;
    float cy[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = y0 + r - h;
        cy[r] = corr[y < 0 ? -y : y];
    }
    const float* zf = reinterpret_cast<const float*>(zb) + tid + 2 * (tid >> 6);
    float cv[NC];
    static_for<NC>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if constexpr ((j + 1) * NTB <= h)
            cv[j] = corr[(h - j * NTB) - tid];   // |x| = h - col
        else
            cv[j] = corr[tid + (j * NTB - h)];   // |x| = col - h
    });
    static_for<NC>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int C = ((j + 1) * NTB <= h) ? j * NTB + (G - h) : j * NTB - h;
        constexpr int off = C + C / 32;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const float val = fabsf(zf[r * L2 + off]) * (cv[j] * cy[r]);
            __builtin_nontemporal_store(val, out0 + (size_t)r * I + tid + j * NTB);
            lmax = fmaxf(lmax, val);
        }
    });
    return lmax;
}

template <int M, int ROWS>
__global__ __launch_bounds__(FftPlan<M>::nt * ROWS) void k_stage2(const float2* __restrict__ kt,
                                                                   const SnapMeta* __restrict__ meta,
                                                                   unsigned* __restrict__ pmax,
                                                                   const float2* __restrict__ tw,
                                                                   const float2* __restrict__ wpost,
                                                                   const float* __restrict__ corr,
                                                                   float* __restrict__ pre, int I)
{
    constexpr int NT = FftPlan<M>::nt;
    constexpr int L = lds_len(M + 1);
    __shared__ float2 zb[ROWS * L];
    __shared__ unsigned shm[NT * ROWS / 64];
    const int s = blockIdx.y;
    const int y0 = xcd_remap(blockIdx.x, gridDim.x) * ROWS;
    const int grp = threadIdx.x / NT, t = threadIdx.x % NT;
    const int cmax = min(meta[s].rmax, M);
    if (cmax < M / 2) {
        C2RGatherLow<M, ROWS> gth;
        gth.load(c2r_rows<M>(kt, s, y0, I), wpost, cmax, threadIdx.x);
        gth.store(zb, cmax, threadIdx.x);
    } else {
        C2RGather<M, ROWS> gth;
        gth.load(c2r_rows<M>(kt, s, y0, I), wpost, cmax, threadIdx.x);
        gth.store(zb, threadIdx.x);
    }
    __syncthreads();
    float2* z = zb + grp * L;
    fft_inverse<M>(z, tw, t);
    float* const out0 = pre + (size_t)s * I * I + (size_t)y0 * I;
    float lmax;
    if constexpr (c2r_fixed_ok<M, NT * ROWS>()) {
        lmax = I == c2r_image<M>() ? c2r_output_fixed<M, ROWS, NT * ROWS>(zb, corr, out0, y0, threadIdx.x)
                                   : c2r_output_rows<M, ROWS, NT * ROWS>(zb, corr, out0, y0, I, threadIdx.x);
    } else {
        lmax = c2r_output_rows<M, ROWS, NT * ROWS>(zb, corr, out0, y0, I, threadIdx.x);
    }
    const unsigned bm = block_max_u<NT * ROWS>(__float_as_uint(lmax), shm);
    if (threadIdx.x == 0) atomicMax(max_slot(pmax, s, 0, (int)(blockIdx.x % kMaxSlots)), bm);
}

// ----------------------------------------------------------------- splat

__device__ __forceinline__ void sin_shift(const SnapGeom& g, double q1, double q2, double& o1, double& o2)
{
    o1 = q1;
    o2 = q2;
    if (!g.sin_on) return;
    // l = -cell (P1 - c0), m = cell (P2 - c0), c0 = I/2 + 1 (FIPkernels.cu:315-316)
    const float l = (float)(-g.delta * (q1 - 1.0));
    const float mm = (float)(g.delta * (q2 - 1.0));
    const float r2 = l * l + mm * mm;
    if (g.series) {
        // every source of the image has r^2 <= 0.01 (checked in k_reset)
        const float np = r2 * fmaf(r2, fmaf(r2, fmaf(r2, 0.0390625f, 0.0625f), 0.125f), 0.5f);
        o1 = q1 + g.sx * (double)np;
        o2 = q2 + g.sy * (double)np;
    } else if (r2 <= 1.0f) {
        // 1 - sqrt(1 - r^2) = r^2 / (1 + sqrt(1 - r^2)); hardware sqrt/rcp (1 ulp)
        const float np = r2 * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f - r2));
        o1 = q1 + g.sx * (double)np;
        o2 = q2 + g.sy * (double)np;
    }
}

// a * b + c for a, b < 2^24 in one full-rate instruction (the compiler
// otherwise forms a quarter-rate 32-bit multiply).
__device__ __forceinline__ unsigned mad_u24(unsigned a, unsigned b, unsigned c)
{
    unsigned r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// a * b + c for |a|, |b| < 2^23 (signed 24-bit) in one full-rate instruction;
// b is uniform (a scalar register).
__device__ __forceinline__ int mad_i24(int a, int b, int c)
{
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// FP32 copy of the geometry the splat needs per source.
struct GeomF {
    float m00, m10;   // affine step per source column
    float sx, sy, delta;
    int sin_on, series;
};
__device__ __forceinline__ GeomF geom_f(const SnapGeom& g)
{
    GeomF f;
    f.m00 = (float)g.m00;
    f.m10 = (float)g.m10;
    f.sx = (float)g.sx;
    f.sy = (float)g.sy;
    f.delta = (float)g.delta;
    f.sin_on = g.sin_on;
    f.series = g.series;
    return f;
}

// Exact integer accumulation: contributions (non-negative) are scaled by a
// power of two, rounded half up to integers and summed as int32; the scale
// keeps every pixel sum below 2^29 (plus half a unit per contribution), so
// each sum is exact and independent of the order of the LDS atomics.
__device__ __forceinline__ void acc_exact(int* a, float x)
{
    atomicAdd(a, cvt_round_half_up_i32(x));
}

// Non-temporal 16-byte load for the trigger, which reads each image byte
// once.  Measured on the one-window trigger: 157 -> 128 us at 8192^2 (three
// images of 268 MB), unchanged at 4096^2; same bits.
typedef float fip_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float* p)
{
    const fip_f4 v = __builtin_nontemporal_load(reinterpret_cast<const fip_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Trigger terms of one pixel (FIPkernels.cu:499-552): dd = ||a - b| - |b - c||,
// its sum and maximum, and the sum of the guarded ratio min(dd / b', 1) with
// b' = b, or 1e-6 where b <= 0 (FIPkernels.cu:811).  Every trigger form adds
// a lane's pixels in the same order through this function, so their sums
// agree bit for bit.
__device__ __forceinline__ void tlisi_px(float a, float b, float c, float& sum, float& mx, float& sr)
{
    const float dd = fabsf(fabsf(a - b) - fabsf(b - c));
    const float b2 = (b <= 0.0f) ? 1e-6f : b;
    sum += dd;
    mx = fmaxf(mx, dd);
    const float ratio = dd / b2;
    sr += (ratio < 1.0f) ? ratio : 1.0f;
}

// tLISI of a T x T tile from its reduced terms and the window's maximum m3
// (FIPkernels.cu:550).
__device__ __forceinline__ float tlisi_value(float sum, float mx, float sr, float m3, int T)
{
    const float Tf = (float)T;
    return 1.0f - (sum / Tf / Tf) * mx * (sr / Tf / Tf) / m3 / m3;
}

// Per target tile: bounding box and per-row ranges of the source pixels whose
// bilinear footprint can reach the tile, then the forward splat
// (finalinterp, FIPkernels.cu:378-413) accumulated in an LDS tile with a
// one-pixel apron.  The candidate sources of a batch of rows are numbered
// row by row; lane work is a contiguous range of that numbering, so every
// lane of every wave has a source whatever the rotation of the map.  A
// lane finds its row from a table of the rows in which each run of 64
// numbers starts (usually no further step) and keeps four independent
// loads and position chains in flight; at most 64 VGPRs keep eight waves
// per SIMD resident.  Snapshot s0 + blockIdx.y.
//
// TRIG: the last snapshot of a three-snapshot window with the trigger fused
// in.  Its image is stored only if store_img; each wave takes one 32 x 32
// trigger tile of the 64 x 64 target tile, reads the window's first two
// images (img_a, img_b, written by an earlier launch) and this tile's values
// from LDS, and stores the tile's reduced terms (sum, max, ratio sum) to part;
// the tLISI follows once every image maximum is known (k_trigger_finish).
template <int TS, bool TRIG>
__global__ __launch_bounds__(256, 8) void k_splat(const float* __restrict__ pre, const SnapMeta* __restrict__ meta,
                                               const SnapGeom* __restrict__ geom, unsigned* __restrict__ pmax,
                                               float* __restrict__ img, int I, int s0,
                                               const float* __restrict__ img_a, const float* __restrict__ img_b,
                                               float* __restrict__ part, int store_img)
{
    constexpr int NT = 256;
    constexpr int NW = NT / 64;           // waves
    constexpr int ROW_CAP = 2 * TS + 32;   // source rows handled per pass
    constexpr int A = TS + 2;             // apron: targets -1 .. TS
    constexpr int GN = 8;                 // consecutive sources of a row per lane
    constexpr int CHK = 512;              // runs of 64 numbered groups per window
    static_assert(GN % 4 == 0 && (GN & (GN - 1)) == 0, "groups of 16-byte loads");
    static_assert((A * A) % 4 == 0, "16-byte clears");
    __shared__ __align__(16) int acc[A * A];
    __shared__ int rlo[ROW_CAP], rcnt[ROW_CAP], rofs[ROW_CAP + 1];
    __shared__ unsigned char crow[CHK];   // row in which each run of 64 starts
    __shared__ double qb[6];
    __shared__ int bb[4];
    __shared__ unsigned shm[NT / 64];
    __shared__ unsigned smaxp;
    const int s = s0 + (int)blockIdx.y;
    const int ntx = I / TS;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int R0 = (tile / ntx) * TS, C0 = (tile % ntx) * TS;
    const SnapGeom g = geom[s];
    const SnapMeta m = meta[s];
    const int h = I / 2;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) {
        int i0 = 0, i1 = I - 1, j0 = 0, j1 = I - 1;
        // q-space (affine position - h) box that can reach the open target box
        // (C0-1, C0+TS) x (R0-1, R0+TS): the target box shifted by the SIN
        // shift at the tile centre, widened by its variation over the tile.
        double ql1 = -1e300, qh1 = 1e300, ql2 = -1e300, qh2 = 1e300;
        if (g.curv >= 0.0) {
            const double t1 = (double)C0 + 0.5 * (TS - 1) - h, t2 = (double)R0 + 0.5 * (TS - 1) - h;
            double q1 = t1, q2 = t2;
            // At most four fixed-point steps; the residual below bounds what is
            // left, so stopping once a step moves less than 1e-4 px keeps the
            // box conservative.
            for (int it = 0; it < 4; ++it) {
                double o1, o2;
                sin_shift(g, q1, q2, o1, o2);
                const double n1 = t1 - (o1 - q1), n2 = t2 - (o2 - q2);
                const bool done = fabs(n1 - q1) + fabs(n2 - q2) < 1e-4;
                q1 = n1;
                q2 = n2;
                if (done) break;
            }
            double o1, o2;
            sin_shift(g, q1, q2, o1, o2);
            const double d1 = o1 - q1, d2 = o2 - q2;
            // Any q with q + d(q) in the box lies within (halfdiag + residual)/(1 - kappa)
            // of (q1, q2), so |d(q) - d| <= kappa times that.
            const double res = fabs(o1 - t1) + fabs(o2 - t2);
            const double halfdiag = 0.7071067811865476 * (TS + 1);
            const double dv = g.kappa * (halfdiag + res) / (1.0 - g.kappa) + 0.02;
            ql1 = (double)(C0 - 1) - h - d1 - dv;
            qh1 = (double)(C0 + TS) - h - d1 + dv;
            ql2 = (double)(R0 - 1) - h - d2 - dv;
            qh2 = (double)(R0 + TS) - h - d2 + dv;
            double mini = 1e300, maxi = -1e300, minj = 1e300, maxj = -1e300;
            for (int k = 0; k < 4; ++k) {
                const double c1 = (k & 1) ? qh1 : ql1, c2 = (k & 2) ? qh2 : ql2;
                const double a = g.i00 * c1 + g.i01 * c2 + h, b = g.i10 * c1 + g.i11 * c2 + h;
                mini = fmin(mini, a);
                maxi = fmax(maxi, a);
                minj = fmin(minj, b);
                maxj = fmax(maxj, b);
            }
            // clamp before converting (degenerate V can give huge values)
            mini = fmin(fmax(mini, -4.0), (double)I + 4.0);
            maxi = fmin(fmax(maxi, -4.0), (double)I + 4.0);
            minj = fmin(fmax(minj, -4.0), (double)I + 4.0);
            maxj = fmin(fmax(maxj, -4.0), (double)I + 4.0);
            i0 = max(0, (int)floor(mini) - 1);
            i1 = min(I - 1, (int)ceil(maxi) + 1);
            j0 = max(0, (int)floor(minj) - 1);
            j1 = min(I - 1, (int)ceil(maxj) + 1);
        }
        qb[0] = ql1;
        qb[1] = qh1;
        qb[2] = ql2;
        qb[3] = qh2;
        qb[4] = g.im00;
        qb[5] = g.im10;
        bb[0] = i0;
        bb[1] = i1;
        bb[2] = j0;
        bb[3] = j1;
    }
    {
        int4* a4 = reinterpret_cast<int4*>(acc);
        for (int q = tid; q < A * A / 4; q += NT) a4[q] = make_int4(0, 0, 0, 0);
    }
    if (wid == 1) {
        // Max of the corrected image from its slots (one per lane), while
        // thread 0 computes the box.
        const unsigned v = wave_max_u(*max_slot(pmax, s, 0, lane));
        if (lane == 0) smaxp = v;
    }
    __syncthreads();
    const int i0 = bb[0], i1 = bb[1], j0 = bb[2], j1 = bb[3];
    // Accepted target offsets of a footprint's top-left corner: the apron
    // -1..TS-1 intersected with |LL| < h - 1, i.e. floor(LL) in [1-h, h-2]
    // (floor(LL) >= 2-h when LL is an integer).
    const int cxlo = max(-1, 1 - C0), cxn = min(TS - 1, I - 2 - C0) - cxlo;
    const int rylo = max(-1, 1 - R0), ryn = min(TS - 1, I - 2 - R0) - rylo;
    const bool cedge = 1 - C0 >= -1, redge = 1 - R0 >= -1;   // tile on the left / top image edge
    // Without a curvature bound positions are not bounded a priori.
    const bool unbounded = g.curv < 0.0;
    const int H = j1 - j0 + 1;
    // A pixel sums at most ccap contributions of at most max(pre)/N each
    // (1/N of FIPkernels.cu:391); 2^k * ccap * max(pre)/N < 2^29.
    int e = 0;
    const unsigned maxp_bits = smaxp;
    frexpf(fmaxf((float)g.ccap * (__uint_as_float(maxp_bits) / (float)m.norm_n), 1e-30f), &e);
    const int ks = min(29 - e, 120);
    const float vscale = ldexpf(1.0f, ks) / (float)m.norm_n;
    const float iscale = ldexpf(1.0f, -ks);
    const GeomF gf = geom_f(g);
    const float* ps = pre + (size_t)s * I * I;
    for (int rb = 0; rb < H; rb += ROW_CAP) {
        const int nrow = min(ROW_CAP, H - rb);
        // Per source row: the i range whose affine position lies in the q box.
        for (int r = tid; r < nrow; r += NT) {
            const double b = (double)(j0 + rb + r - h);
            double alo = -1e300, ahi = 1e300;
            const double mA[2] = {g.m00, g.m10}, mB[2] = {g.m01, g.m11}, iA[2] = {qb[4], qb[5]};
            // The q box is read from LDS here rather than held in registers
            // across the source loop (eight VGPRs; holding them spilled).
            const double lo[2] = {qb[0], qb[2]}, hi[2] = {qb[1], qb[3]};
            for (int k = 0; k < 2; ++k) {
                const double cst = mB[k] * b;
                if (fabs(mA[k]) > 1e-12) {
                    // The range is widened by one pixel below, so a rounded
                    // reciprocal is as good as a division here.
                    double x = (lo[k] - cst) * iA[k], y = (hi[k] - cst) * iA[k];
                    if (x > y) {
                        const double t = x;
                        x = y;
                        y = t;
                    }
                    alo = fmax(alo, x);
                    ahi = fmin(ahi, y);
                } else if (cst < lo[k] || cst > hi[k]) {
                    alo = 1e300;
                    ahi = -1e300;
                }
            }
            alo = fmax(alo, (double)(i0 - h - 2));
            ahi = fmin(ahi, (double)(i1 - h + 2));
            int ia = i0, n = 0;
            if (alo <= ahi) {
                ia = max(i0, (int)floor(alo + h) - 1);
                n = max(0, min(i1, (int)ceil(ahi + h) + 1) - ia + 1);
            }
            // Groups of GN pixels aligned to i % GN == 0 that cover [ia, ia + n).
            rlo[r] = ia / GN;
            rcnt[r] = n > 0 ? ((ia + n - 1) / GN) - (ia / GN) + 1 : 0;
        }
        __syncthreads();
        // Exclusive prefix of the per-row counts (wave 0).
        if (wid == 0) {
            int run = 0;
            for (int b0 = 0; b0 < nrow; b0 += 64) {
                const int r = b0 + lane;
                const int v = r < nrow ? rcnt[r] : 0;
                int x = v;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int y = __shfl_up(x, d);
                    if (lane >= d) x += y;
                }
                if (r < nrow) rofs[r] = run + x - v;
                run += __shfl(x, 63);
            }
            if (lane == 0) rofs[nrow] = run;
        }
        __syncthreads();
        const int total = rofs[nrow];
        for (int f0 = 0; f0 < total; f0 += CHK * 64) {
            const int nf = min(CHK * 64, total - f0);
            // Row of the first number of every run of 64 in [f0, f0 + nf).
            for (int r = tid; r < nrow; r += NT) {
                const int a = max(rofs[r], f0), b = min(rofs[r] + rcnt[r], f0 + nf);
                for (int k = (a - f0 + 63) >> 6; (k << 6) < b - f0; ++k) crow[k] = (unsigned char)r;
            }
            __syncthreads();
            // The source loop, specialised on the tile: EDGE for tiles on the
            // left or top image edge (an integer position on the bound is
            // outside there), UNB for maps without a curvature bound
            // (positions checked for size).  Interior tiles of bounded maps,
            // nearly all of them, take neither check.
            auto sources = [&](auto EDGE, auto UNB, auto SER) {
                constexpr bool edge = decltype(EDGE)::value, unb = decltype(UNB)::value;
                // SIN shift: 1 by its series, 0 by its closed form (a map
                // without one has sx = sy = 0, which adds exact zeros), -1 as
                // the geometry's flags say
                constexpr int ser = decltype(SER)::value;
                for (int base = 0; base < nf; base += NT) {
                    const int fl = base + tid;
                    if (fl >= nf) break;
                    const int f = f0 + fl;
                    int r = crow[fl >> 6];
                    while (rofs[r + 1] <= f) ++r;
                    const int gi = rlo[r] + (f - rofs[r]);
                    const int jj = j0 + rb + r;
                    // Aligned 16-byte loads (I % GN == 0, so a group never
                    // crosses the end of a row).
                    float pv[GN];
                    const float4* src = reinterpret_cast<const float4*>(ps + (size_t)jj * I + GN * gi);
#pragma unroll
                    for (int u = 0; u < GN / 4; ++u) {
                        const float4 q = src[u];
                        pv[4 * u] = q.x;
                        pv[4 * u + 1] = q.y;
                        pv[4 * u + 2] = q.z;
                        pv[4 * u + 3] = q.w;
                    }
                    // Group origin in FP64: the affine position of its first
                    // pixel (clamped: positions beyond 1e6 px are rejected
                    // below), split into an integer part O and a fraction.  The
                    // pixels are placed in FP32 relative to O (local
                    // coordinates stay within a few pixels, ~1e-7 px), and only
                    // integer parts are recombined.
                    const double a0 = (double)(GN * gi - h), b = (double)(jj - h);
                    double q1d = fma(g.m00, a0, g.m01 * b), q2d = fma(g.m10, a0, g.m11 * b);
                    if constexpr (unb) {
                        // bounded maps keep positions within the image's reach
                        q1d = fmin(fmax(q1d, -1e6), 1e6);
                        q2d = fmin(fmax(q2d, -1e6), 1e6);
                    }
                    const double O1 = floor(q1d), O2 = floor(q2d);
                    const float r1 = (float)(q1d - O1), r2 = (float)(q2d - O2);
                    // target offsets of the origin in the tile, relative to the
                    // accepted ranges' starts, and the LDS byte address of the
                    // origin's apron cell
                    const int c1 = (int)O1 + h - C0, c2 = (int)O2 + h - R0;
                    const int e1 = c1 - cxlo, e2 = c2 - rylo;
                    unsigned gb = 4u * (unsigned)(__mul24(c2 + 1, A) + c1 + 1);
                    asm volatile("" : "+v"(gb));   // kept whole, so gb + 4 f1 is one instruction
                    // SIN direction cosines along the group: l = -cell (P1 - c0),
                    // m = cell (P2 - c0), c0 = I/2 + 1 (FIPkernels.cu:315-316),
                    // rounded once from FP64.
                    const float l0 = (float)fma(-g.delta, q1d, g.delta), dl = -gf.delta * gf.m00;
                    const float n0 = (float)fma(g.delta, q2d, -g.delta), dn = gf.delta * gf.m10;
                    // Series form: source k of the group lands at
                    // (Q1 + B1 k + C1 k^2, Q2 + B2 k + C2 k^2) relative to the
                    // origin, the affine step plus the quadratic Taylor model in
                    // k of the SIN shift (sx, sy) n'(k) (its remainder is
                    // bounded in k_reset).  With r^2(k) = R0 + R1 k + R2 k^2 and
                    // the series f(x) = x/2 + x^2/8 + x^3/16 + 5x^4/128,
                    // n'(k) ~ f(R0) + f'(R0) R1 k + (f'(R0) R2 + f''(R0) R1^2 / 2) k^2.
                    float Q1 = r1, B1 = gf.m00, C1 = 0.0f, Q2 = r2, B2 = gf.m10, C2 = 0.0f;
                    if constexpr (ser == 1) {
                        const float R0 = fmaf(l0, l0, n0 * n0);
                        const float R1 = fmaf(l0, 2.0f * dl, n0 * (2.0f * dn));
                        const float R2 = fmaf(dl, dl, dn * dn);
                        const float f0 = R0 * fmaf(R0, fmaf(R0, fmaf(R0, 0.0390625f, 0.0625f), 0.125f), 0.5f);
                        const float f1 = fmaf(R0, fmaf(R0, fmaf(R0, 0.15625f, 0.1875f), 0.25f), 0.5f);
                        const float f2h = fmaf(R0, fmaf(R0, 0.234375f, 0.1875f), 0.125f);
                        const float b = f1 * R1, c = fmaf(f1, R2, f2h * R1 * R1);
                        Q1 = fmaf(gf.sx, f0, r1);
                        Q2 = fmaf(gf.sy, f0, r2);
                        B1 = fmaf(gf.sx, b, gf.m00);
                        B2 = fmaf(gf.sy, b, gf.m10);
                        C1 = gf.sx * c;
                        C2 = gf.sy * c;
                    }
#pragma unroll
                    for (int k = 0; k < GN; ++k) {
                        float o1, o2;
                        if constexpr (ser == 1) {
                            o1 = k == 0 ? Q1 : fmaf((float)(k * k), C1, fmaf((float)k, B1, Q1));
                            o2 = k == 0 ? Q2 : fmaf((float)(k * k), C2, fmaf((float)k, B2, Q2));
                        } else {
                            o1 = fmaf((float)k, gf.m00, r1);
                            o2 = fmaf((float)k, gf.m10, r2);
                            if (ser == 0 || gf.sin_on) {
                                const float l = fmaf((float)k, dl, l0), mm = fmaf((float)k, dn, n0);
                                const float rr = l * l + mm * mm;
                                float np = 0.0f;
                                if (ser < 0 && gf.series) {
                                    // every source of the image has r^2 <= 0.01 (checked in k_reset)
                                    np = rr * fmaf(rr, fmaf(rr, fmaf(rr, 0.0390625f, 0.0625f), 0.125f), 0.5f);
                                } else if (rr <= 1.0f) {
                                    // 1 - sqrt(1 - r^2) = r^2 / (1 + sqrt(1 - r^2)); hardware sqrt/rcp
                                    np = rr * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f - rr));
                                }
                                o1 = fmaf(gf.sx, np, o1);
                                o2 = fmaf(gf.sy, np, o2);
                            }
                        }
                        bool fin = true;
                        if constexpr (unb) {
                            // finite and small enough for the integer recombination
                            fin = fabsf(o1) < 1048576.0f && fabsf(o2) < 1048576.0f;
                            o1 = fin ? o1 : 0.0f;
                            o2 = fin ? o2 : 0.0f;
                        }
                        // floor as an integer and the fraction o - floor(o) in
                        // [0, 1): the bilinear weights of the two axes
                        const int f1 = cvt_floor_i32(o1), f2 = cvt_floor_i32(o2);
                        const float wl1 = __builtin_amdgcn_fractf(o1), wm1 = __builtin_amdgcn_fractf(o2);
                        // Tile apron and |LL|, |MM| < h - 1 (FIPkernels.cu:389) as
                        // one unsigned range per axis; on the left / top image
                        // edge an integer position on the bound is outside.
                        const int u1 = f1 + e1, u2 = f2 + e2;
                        bool ok = (unsigned)u1 <= (unsigned)cxn && (unsigned)u2 <= (unsigned)ryn;
                        if constexpr (unb) ok = ok && fin;
                        if constexpr (edge)
                            ok = ok && !(cedge && u1 == 0 && wl1 == 0.0f) && !(redge && u2 == 0 && wm1 == 0.0f);
                        if (ok) {
                            // FIPkernels.cu:395-410: (fM,fL), (cM,fL), (fM,cL), (cM,cL)
                            // with weights (1-a)(1-b), (1-a)b, a(1-b), ab; each
                            // product is split off its row by one subtraction.  A
                            // zero weight on an integer position adds an exact zero.
                            const float vs = pv[k] * vscale;
                            const float t1 = wl1 * vs, t0 = vs - t1;
                            const float p01 = t0 * wm1, p11 = t1 * wm1;
                            const float p00 = t0 - p01, p10 = t1 - p11;
                            int* p0 = reinterpret_cast<int*>(reinterpret_cast<char*>(acc) +
                                                             (unsigned)mad_i24(f2, 4 * A, (int)gb + 4 * f1));
                            acc_exact(p0, p00);
                            acc_exact(p0 + A, p01);
                            acc_exact(p0 + 1, p10);
                            acc_exact(p0 + A + 1, p11);
                        }
                    }
                }
            };
            // Bounded maps take the SIN form as a compile-time choice: 2-5%
            // faster than the per-source test of the flags.
            using Series = std::integral_constant<int, 1>;
            using Closed = std::integral_constant<int, 0>;
            if (unbounded) {
                sources(std::true_type{}, std::true_type{}, std::integral_constant<int, -1>{});
            } else if (cedge || redge) {
                if (gf.series)
                    sources(std::true_type{}, std::false_type{}, Series{});
                else
                    sources(std::true_type{}, std::false_type{}, Closed{});
            } else {
                if (gf.series)
                    sources(std::false_type{}, std::false_type{}, Series{});
                else
                    sources(std::false_type{}, std::false_type{}, Closed{});
            }
            __syncthreads();   // crow is rewritten by the next window
        }
        __syncthreads();   // rlo/rcnt are rewritten by the next row batch
    }
    float* out = img + (size_t)s * I * I;
    float lmax = 0.0f;
    // Four consecutive pixels per thread per pass, one 16-byte store; each
    // pass covers NT / (TS / 4) rows, and the pointers step by that.
    {
        constexpr int RSTEP = NT / (TS / 4);
        static_assert((TS * TS / 4) % NT == 0, "whole passes");
        const int rr = tid / (TS / 4), cc = 4 * (tid % (TS / 4));
        const int* a = acc + (rr + 1) * A + cc + 1;
        float* o = out + (size_t)(R0 + rr) * I + C0 + cc;
#pragma unroll
        for (int q = 0; q < TS * TS / 4 / NT; ++q) {
            const float4 v = make_float4((float)a[0] * iscale, (float)a[1] * iscale, (float)a[2] * iscale,
                                         (float)a[3] * iscale);
            if (!TRIG || store_img) *reinterpret_cast<float4*>(o) = v;
            lmax = fmaxf(lmax, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
            a += RSTEP * A;
            o += (size_t)RSTEP * I;
        }
    }
    const unsigned bmx = block_max_u<NT>(__float_as_uint(lmax), shm);
    if (tid == 0) atomicMax(max_slot(pmax, s, 1, (int)(blockIdx.x % kMaxSlots)), bmx);
    if constexpr (TRIG) {
        // The trigger tile of this wave; lanes take the pixels, and add them
        // in the order, of k_trigger (16-byte groups lane + 64 u), so the
        // reduced terms are the same bits as k_trigger's.
        static_assert(TS == 64, "four 32 x 32 trigger tiles per workgroup, one per wave");
        constexpr int T = 32;
        const int tr = R0 + T * (wid >> 1), tc = C0 + T * (wid & 1);
        float sum = 0.0f, mx = 0.0f, sr = 0.0f;
        float4 av[4], bv[4], cv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = lane + 64 * u, row = p / (T / 4), col = 4 * (p % (T / 4));
            const size_t q = (size_t)(tr + row) * I + (size_t)(tc + col);
            av[u] = load_stream(img_a + q);
            bv[u] = load_stream(img_b + q);
            // the values this launch would have stored
            const int* ac = acc + (tr - R0 + row + 1) * A + (tc - C0 + col + 1);
            cv[u] = make_float4((float)ac[0] * iscale, (float)ac[1] * iscale, (float)ac[2] * iscale,
                                (float)ac[3] * iscale);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tlisi_px(av[u].x, bv[u].x, cv[u].x, sum, mx, sr);
            tlisi_px(av[u].y, bv[u].y, cv[u].y, sum, mx, sr);
            tlisi_px(av[u].z, bv[u].z, cv[u].z, sum, mx, sr);
            tlisi_px(av[u].w, bv[u].w, cv[u].w, sum, mx, sr);
        }
        sum = wave_sum_f(sum);
        mx = wave_max_f(mx);
        sr = wave_sum_f(sr);
        if (lane == 0) {
            float* pp = part + 3 * ((size_t)(tr / T) * (I / T) + (size_t)(tc / T));
            pp[0] = sum;
            pp[1] = mx;
            pp[2] = sr;
        }
    }
}

// tLISI of one window from the fused splat's reduced terms: each wave reduces
// the window's three image maxima from their spread slots (as k_trigger does),
// then each thread finishes one tile.
__global__ __launch_bounds__(256) void k_trigger_finish(const float* __restrict__ part, SnapMeta* __restrict__ meta,
                                                        unsigned* __restrict__ pmax, int s1, int s2, int s3, int I,
                                                        int T, float* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const unsigned m1b = wave_max_u(*max_slot(pmax, s1, 1, lane)), m2b = wave_max_u(*max_slot(pmax, s2, 1, lane)),
                   m3b = wave_max_u(*max_slot(pmax, s3, 1, lane));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the reduced maxima for host readers
        meta[s1].maxd_bits = m1b;
        meta[s2].maxd_bits = m2b;
        meta[s3].maxd_bits = m3b;
    }
    const int un = I / T;
    const int tile = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (tile >= un * un) return;
    const float* pp = part + 3 * (size_t)tile;
    const float m3 = fmaxf(__uint_as_float(m1b), fmaxf(__uint_as_float(m2b), __uint_as_float(m3b)));
    out[tile] = tlisi_value(pp[0], pp[1], pp[2], m3, T);
}

// --------------------------------------------------------------- trigger

// One wave per tile (four tiles per workgroup): per-lane partial sums over
// 16-byte loads, then wave reductions only (FIPkernels.cu:499-552).
__global__ __launch_bounds__(256) void k_trigger(const float* __restrict__ img, SnapMeta* __restrict__ meta,
                                                 unsigned* __restrict__ pmax, const int* __restrict__ win, int I,
                                                 int T, float* __restrict__ out)
{
    const int w = blockIdx.y;
    const int un = I / T;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= un * un) return;   // whole wave
    const int ti = tile / un, tj = tile % un;
    const int s1 = win[3 * w], s2 = win[3 * w + 1], s3 = win[3 * w + 2];
    // Image maxima from the spread slots (one slot per lane).
    const unsigned q1 = *max_slot(pmax, s1, 1, lane), q2 = *max_slot(pmax, s2, 1, lane),
                   q3 = *max_slot(pmax, s3, 1, lane);
    const float* d1 = img + (size_t)s1 * I * I;
    const float* d2 = img + (size_t)s2 * I * I;
    const float* d3 = img + (size_t)s3 * I * I;
    float sum = 0.0f, mx = 0.0f, sr = 0.0f;
    auto px = [&](float a, float b, float c) { tlisi_px(a, b, c, sum, mx, sr); };
    if ((T & 3) == 0) {
        // 16-byte loads: rows of a tile start at multiples of 4 floats; the
        // loads of four passes are issued together.
        const int T4 = T >> 2, n4 = T * T4;
        for (int p0 = lane; p0 < n4; p0 += 4 * 64) {
            float4 a[4], b[4], c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int p = p0 + u * 64;
                if (p < n4) {
                    const size_t q = (size_t)(ti * T + p / T4) * I + (size_t)(tj * T + 4 * (p % T4));
                    a[u] = load_stream(d1 + q);
                    b[u] = load_stream(d2 + q);
                    c[u] = load_stream(d3 + q);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (p0 + u * 64 < n4) {
                    px(a[u].x, b[u].x, c[u].x);
                    px(a[u].y, b[u].y, c[u].y);
                    px(a[u].z, b[u].z, c[u].z);
                    px(a[u].w, b[u].w, c[u].w);
                }
            }
        }
    } else {
        for (int p = lane; p < T * T; p += 64) {
            const size_t q = (size_t)(ti * T + p / T) * I + (size_t)(tj * T + p % T);
            px(d1[q], d2[q], d3[q]);
        }
    }
    sum = wave_sum_f(sum);
    mx = wave_max_f(mx);
    sr = wave_sum_f(sr);
    const unsigned m1b = wave_max_u(q1), m2b = wave_max_u(q2), m3b = wave_max_u(q3);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the reduced maxima for host readers (every window writes the same values)
        meta[s1].maxd_bits = m1b;
        meta[s2].maxd_bits = m2b;
        meta[s3].maxd_bits = m3b;
    }
    if (lane == 0) {
        const float m3 = fmaxf(__uint_as_float(m1b), fmaxf(__uint_as_float(m2b), __uint_as_float(m3b)));
        out[(size_t)w * un * un + tile] = tlisi_value(sum, mx, sr, m3, T);
    }
}

// Consecutive windows (s0+w, s0+w+1, s0+w+2), w < nwin, of a batch: one wave
// per tile handles up to kRunWin consecutive windows, so each image is read
// once per kRunWin windows instead of three times per window.
constexpr int kRunWin = 8;

__global__ __launch_bounds__(256) void k_trigger_run(const float* __restrict__ img, SnapMeta* __restrict__ meta,
                                                     unsigned* __restrict__ pmax, int s0, int nwin, int I, int T,
                                                     float* __restrict__ out)
{
    constexpr int K = kRunWin, NI = kRunWin + 2;
    const int un = I / T;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= un * un) return;   // whole wave
    const int w0 = blockIdx.y * K, nk = min(K, nwin - w0), ni = nk + 2;
    const int i0 = s0 + w0;
    const int ti = tile / un, tj = tile % un;
    // Image maxima from the spread slots (one slot per lane).
    unsigned md[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) md[j] = wave_max_u(j < ni ? *max_slot(pmax, i0 + j, 1, lane) : 0u);
    float sum[K], mx[K], sr[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sum[k] = mx[k] = sr[k] = 0.0f;
    auto px = [&](int k, float a, float b, float c) { tlisi_px(a, b, c, sum[k], mx[k], sr[k]); };
    const size_t plane = (size_t)I * I;
    const float* base = img + (size_t)i0 * plane;
    if ((T & 3) == 0) {
        const int T4 = T >> 2, n4 = T * T4;
        for (int p = lane; p < n4; p += 64) {
            const size_t q = (size_t)(ti * T + p / T4) * I + (size_t)(tj * T + 4 * (p % T4));
            float4 d[NI];
#pragma unroll
            for (int j = 0; j < NI; ++j)
                d[j] = j < ni ? load_stream(base + (size_t)j * plane + q)
                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (k < nk) {
                    px(k, d[k].x, d[k + 1].x, d[k + 2].x);
                    px(k, d[k].y, d[k + 1].y, d[k + 2].y);
                    px(k, d[k].z, d[k + 1].z, d[k + 2].z);
                    px(k, d[k].w, d[k + 1].w, d[k + 2].w);
                }
            }
        }
    } else {
        for (int p = lane; p < T * T; p += 64) {
            const size_t q = (size_t)(ti * T + p / T) * I + (size_t)(tj * T + p % T);
            float d[NI];
#pragma unroll
            for (int j = 0; j < NI; ++j) d[j] = j < ni ? base[(size_t)j * plane + q] : 0.0f;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < nk) px(k, d[k], d[k + 1], d[k + 2]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the reduced maxima for host readers
        for (int j = 0; j < ni; ++j) meta[i0 + j].maxd_bits = md[j];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float sk = wave_sum_f(sum[k]), mk = wave_max_f(mx[k]), rk = wave_sum_f(sr[k]);
        if (lane == 0 && k < nk) {
            const float m3 = fmaxf(__uint_as_float(md[k]), fmaxf(__uint_as_float(md[k + 1]), __uint_as_float(md[k + 2])));
            out[(size_t)(w0 + k) * un * un + tile] = tlisi_value(sk, mk, rk, m3, T);
        }
    }
}

// ------------------------------------------------------------- launchers

const char* stage_name(int st)
{
    static const char* names[kStCount] = {"reset",   "bin",   "scan",          "scatter", "grid_colfft",
                                          "row_c2r", "splat", "splat_trigger", "trigger"};
    return (st >= 0 && st < kStCount) ? names[st] : "?";
}

hipError_t launch_reset(const InputView& in, const WorkView& w, const PlanTables& p, int S, float cell,
                        const int* norm_n_dev, int norm_n_const, hipStream_t st)
{
    const int total = max(S * p.G, S * 2 * kMaxSlots);
    hipLaunchKernelGGL(k_reset, dim3((total + 255) / 256), dim3(256), 0, st, in.vin, in.nvis, in.n_const,
                       norm_n_dev, norm_n_const, w.meta, w.geom, w.hist, w.work, w.pmax, S, p.G, p.I, cell);
    return hipGetLastError();
}

hipError_t launch_bin(const InputView& in, const WorkView& w, const PlanTables& p, int S, int nmax,
                      double scale, hipStream_t st)
{
    if (nmax <= 0) return hipSuccess;
    dim3 grid((nmax + kBinThreads * kBinVpt - 1) / (kBinThreads * kBinVpt), S);
    const size_t lds = (size_t)p.G * sizeof(int);
    switch (in.vis_type) {
    case kVisF32:
        hipLaunchKernelGGL(k_bin<float>, grid, dim3(kBinThreads), lds, st, in, w.rec, w.bkslot, w.hist, w.meta, p.G, scale);
        break;
    case kVisBF16:
        hipLaunchKernelGGL(k_bin<__hip_bfloat16>, grid, dim3(kBinThreads), lds, st, in, w.rec, w.bkslot, w.hist, w.meta,
                           p.G, scale);
        break;
    case kVisF16:
        hipLaunchKernelGGL(k_bin<__half>, grid, dim3(kBinThreads), lds, st, in, w.rec, w.bkslot, w.hist, w.meta, p.G,
                           scale);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan(const WorkView& w, const PlanTables& p, int S, hipStream_t st)
{
    if (p.G > kMaxGrid) return hipErrorInvalidValue;   // a scan thread holds at most 12 counts
    hipLaunchKernelGGL(k_scan, dim3(S), dim3(1024), 0, st, w.hist, w.bstart, p.G);
    return hipGetLastError();
}

hipError_t launch_scatter(const WorkView& w, const PlanTables& p, int S, int nmax, hipStream_t st)
{
    if (nmax <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((nmax + 255) / 256, S), dim3(256), 0, st, w.rec, w.bkslot, w.bstart,
                       w.sorted, w.meta, w.stride, p.G);
    return hipGetLastError();
}

template <int G, int KR>
static void stage1_dispatch(const WorkView& w, const PlanTables& p, int S, bool persistent, hipStream_t st)
{
    constexpr int NT = stage1_threads<G>();
    if (persistent) {
        int dev = 0, ncu = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stage1_persistent<G, KR>, NT, 0);
        const int nb = max(1, ncu * max(per_cu, 1));
        hipLaunchKernelGGL((k_stage1_persistent<G, KR>), dim3(nb), dim3(NT), 0, st, w.sorted, w.bstart, w.meta,
                           p.tw_g, w.kt, w.stride, p.I, S, w.work);
    } else {
        hipLaunchKernelGGL((k_stage1<G, KR>), dim3((G / 2 + 1) * S), dim3(NT), 0, st, w.sorted, w.bstart, w.meta,
                           p.tw_g, w.kt, w.stride, p.I);
    }
}

// Records per thread held in registers while a column is gridded.
template <int G>
static void stage1_pick(const WorkView& w, const PlanTables& p, int S, const Tuning& tu, hipStream_t st)
{
    if (tu.stage1_records >= 8)
        stage1_dispatch<G, 8>(w, p, S, tu.persistent != 0, st);
    else if (tu.stage1_records <= 2)
        stage1_dispatch<G, 2>(w, p, S, tu.persistent != 0, st);
    else
        stage1_dispatch<G, 4>(w, p, S, tu.persistent != 0, st);
}

hipError_t launch_stage1(const WorkView& w, const PlanTables& p, int S, const Tuning& tu, hipStream_t st)
{
    switch (p.G) {
    case 96: stage1_pick<96>(w, p, S, tu, st); break;
    case 192: stage1_pick<192>(w, p, S, tu, st); break;
    case 384: stage1_pick<384>(w, p, S, tu, st); break;
    case 768: stage1_pick<768>(w, p, S, tu, st); break;
    case 1536: stage1_pick<1536>(w, p, S, tu, st); break;
    case 3072: stage1_pick<3072>(w, p, S, tu, st); break;
    case 6144: stage1_pick<6144>(w, p, S, tu, st); break;
    case 12288: stage1_pick<12288>(w, p, S, tu, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int M, int ROWS>
static void stage2_dispatch(const WorkView& w, const PlanTables& p, int S, hipStream_t st)
{
    constexpr int NT = FftPlan<M>::nt * ROWS;
    hipLaunchKernelGGL((k_stage2<M, ROWS>), dim3(p.I / ROWS, S), dim3(NT), 0, st, w.kt, w.meta, w.pmax, p.tw_m,
                       p.wpost, p.corr, w.pre, p.I);
}

template <int M>
static void stage2_pick(const WorkView& w, const PlanTables& p, int S, const Tuning& tu, hipStream_t st)
{
    // LDS per row is (M + M/32 + 2) * 8 bytes; keep a workgroup under 128 KiB.
    constexpr int row_bytes = lds_len(M + 1) * 8;
    if constexpr (4 * row_bytes <= 128 * 1024 && kKTile % 4 == 0) {
        if (tu.stage2_rows >= 4) {
            stage2_dispatch<M, 4>(w, p, S, st);
            return;
        }
    }
    if constexpr (2 * row_bytes <= 128 * 1024 && kKTile % 2 == 0) {
        if (tu.stage2_rows >= 2) {
            stage2_dispatch<M, 2>(w, p, S, st);
            return;
        }
    }
    stage2_dispatch<M, 1>(w, p, S, st);
}

hipError_t launch_stage2(const WorkView& w, const PlanTables& p, int S, const Tuning& tu, hipStream_t st)
{
    switch (p.M) {
    case 48: stage2_pick<48>(w, p, S, tu, st); break;
    case 96: stage2_pick<96>(w, p, S, tu, st); break;
    case 192: stage2_pick<192>(w, p, S, tu, st); break;
    case 384: stage2_pick<384>(w, p, S, tu, st); break;
    case 768: stage2_pick<768>(w, p, S, tu, st); break;
    case 1536: stage2_pick<1536>(w, p, S, tu, st); break;
    case 3072: stage2_pick<3072>(w, p, S, tu, st); break;
    case 6144: stage2_pick<6144>(w, p, S, tu, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_splat(const WorkView& w, const PlanTables& p, int S, const Tuning& tu, hipStream_t st)
{
    if (S <= 0) return hipSuccess;
    if (tu.splat_tile >= 64 && p.I % 64 == 0) {
        const int nt = (p.I / 64) * (p.I / 64);
        hipLaunchKernelGGL((k_splat<64, false>), dim3(nt, S), dim3(256), 0, st, w.pre, w.meta, w.geom, w.pmax, w.img,
                           p.I, 0, nullptr, nullptr, nullptr, 1);
    } else {
        const int nt = (p.I / 32) * (p.I / 32);
        hipLaunchKernelGGL((k_splat<32, false>), dim3(nt, S), dim3(256), 0, st, w.pre, w.meta, w.geom, w.pmax, w.img,
                           p.I, 0, nullptr, nullptr, nullptr, 1);
    }
    return hipGetLastError();
}

bool splat_trigger_fusable(const PlanTables& p, const Tuning& tu, int tile)
{
    return tu.splat_tile >= 64 && p.I % 64 == 0 && tile == 32;
}

// The reduced trigger terms go to the K tiles' buffer, which the column and
// row passes are done with (3 floats per trigger tile; the buffer holds
// 8 I (g/2 + 1) bytes per snapshot).
static float* trigger_terms(const WorkView& w) { return reinterpret_cast<float*>(w.kt); }

hipError_t launch_splat_trigger(const WorkView& w, const PlanTables& p, int s3, const float* img_a,
                                const float* img_b, bool store, hipStream_t st)
{
    const int nt = (p.I / 64) * (p.I / 64);
    hipLaunchKernelGGL((k_splat<64, true>), dim3(nt, 1), dim3(256), 0, st, w.pre, w.meta, w.geom, w.pmax, w.img, p.I,
                       s3, img_a, img_b, trigger_terms(w), store ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_trigger_finish(const WorkView& w, const PlanTables& p, int s1, int s2, int s3, int tile, float* out,
                                 hipStream_t st)
{
    const int un = p.I / tile;
    hipLaunchKernelGGL(k_trigger_finish, dim3((un * un + 255) / 256), dim3(256), 0, st, trigger_terms(w), w.meta,
                       w.pmax, s1, s2, s3, p.I, tile, out);
    return hipGetLastError();
}

hipError_t launch_trigger(const WorkView& w, const PlanTables& p, const int* win, int nwin, int tile, float* out,
                          hipStream_t st)
{
    const int un = p.I / tile;
    if (un <= 0 || nwin <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trigger, dim3((un * un + 3) / 4, nwin), dim3(256), 0, st, w.img, w.meta, w.pmax, win, p.I,
                       tile, out);
    return hipGetLastError();
}

hipError_t launch_trigger_run(const WorkView& w, const PlanTables& p, int s0, int nwin, int tile, float* out,
                              hipStream_t st)
{
    const int un = p.I / tile;
    if (un <= 0 || nwin <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trigger_run, dim3((un * un + 3) / 4, (nwin + kRunWin - 1) / kRunWin), dim3(256), 0, st,
                       w.img, w.meta, w.pmax, s0, nwin, p.I, tile, out);
    return hipGetLastError();
}

// One wave per snapshot: the slot maxima into meta.
__global__ __launch_bounds__(64) void k_max_finalize(SnapMeta* __restrict__ meta, unsigned* __restrict__ pmax)
{
    const int s = blockIdx.x, lane = threadIdx.x;
    const unsigned mp = wave_max_u(*max_slot(pmax, s, 0, lane));
    const unsigned md = wave_max_u(*max_slot(pmax, s, 1, lane));
    if (lane == 0) {
        meta[s].maxp_bits = mp;
        meta[s].maxd_bits = md;
    }
}

hipError_t launch_max_finalize(const WorkView& w, int S, hipStream_t st)
{
    if (S <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_max_finalize, dim3(S), dim3(64), 0, st, w.meta, w.pmax);
    return hipGetLastError();
}

// ------------------------------------------------------------ FFT check

template <int N>
__global__ __launch_bounds__(FftPlan<N>::nt) void k_fft_check(float2* __restrict__ data, const float2* __restrict__ tw)
{
    constexpr int NT = FftPlan<N>::nt;
    __shared__ float2 fb[lds_len(N)];
    float2* d = data + (size_t)blockIdx.x * N;
    for (int i = threadIdx.x; i < N; i += NT) fb[lds_pad(i)] = d[i];
    __syncthreads();
    fft_inverse<N>(fb, tw, threadIdx.x);
    for (int i = threadIdx.x; i < N; i += NT) d[i] = fb[lds_pad(i)];
}

hipError_t launch_fft_check(int n, float2* data, const float2* tw, int count, hipStream_t st)
{
    switch (n) {
#define FIP_FFT_CASE(N_)                                                                                  \
    case N_:                                                                                              \
        hipLaunchKernelGGL(k_fft_check<N_>, dim3(count), dim3(FftPlan<N_>::nt), 0, st, data, tw);        \
        break;
        FIP_FFT_CASE(48)
        FIP_FFT_CASE(96)
        FIP_FFT_CASE(192)
        FIP_FFT_CASE(384)
        FIP_FFT_CASE(768)
        FIP_FFT_CASE(1536)
        FIP_FFT_CASE(3072)
        FIP_FFT_CASE(6144)
        FIP_FFT_CASE(12288)
#undef FIP_FFT_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ----------------------------------------------------- plan-side helpers

template <int N>
static void twiddles_for(float2* out)
{
    using P = FftPlan<N>;
    int ns = 1, off = 0;
    for (int s = 0; s < P::nstages; ++s) {
        const int R = P::radix[s];
        if (ns > 1) {
            for (int r = 1; r < R; ++r) {
                for (int k = 0; k < ns; ++k) {
                    // exp(+2 pi i r k / (ns R)), argument reduced exactly in integers
                    const long long num = (long long)r * k % ((long long)ns * R);
                    const double ang = 2.0 * M_PI * (double)num / (double)(ns * R);
                    out[off + (r - 1) * ns + k] = make_float2((float)cos(ang), (float)sin(ang));
                }
            }
            off += (R - 1) * ns;
        }
        ns *= R;
    }
}

bool fft_size_supported(int n)
{
    switch (n) {
    case 48: case 96: case 192: case 384: case 768: case 1536: case 3072: case 6144: case 12288:
        return true;
    default:
        return false;
    }
}

void make_stage_twiddles(int n, float2* out)
{
    switch (n) {
    case 48: twiddles_for<48>(out); break;
    case 96: twiddles_for<96>(out); break;
    case 192: twiddles_for<192>(out); break;
    case 384: twiddles_for<384>(out); break;
    case 768: twiddles_for<768>(out); break;
    case 1536: twiddles_for<1536>(out); break;
    case 3072: twiddles_for<3072>(out); break;
    case 6144: twiddles_for<6144>(out); break;
    case 12288: twiddles_for<12288>(out); break;
    default: break;
    }
}

int fft_threads(int n)
{
    switch (n) {
    case 48: return FftPlan<48>::nt;
    case 96: return FftPlan<96>::nt;
    case 192: return FftPlan<192>::nt;
    case 384: return FftPlan<384>::nt;
    case 768: return FftPlan<768>::nt;
    case 1536: return FftPlan<1536>::nt;
    case 3072: return FftPlan<3072>::nt;
    case 6144: return FftPlan<6144>::nt;
    case 12288: return FftPlan<12288>::nt;
    default: return 0;
    }
}

}  // namespace fip

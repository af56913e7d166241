// LDS-resident mixed-radix Stockham inverse FFT for gfx950.
//
// One transform of compile-time length N is held in a padded LDS array and
// processed by NT threads (a "group"); several groups of one workgroup may
// transform independent arrays in lock-step (every stage ends in a workgroup
// barrier).  Each stage loads R elements per butterfly into registers,
// applies the stage twiddles, runs an in-register radix-R butterfly and
// writes back in Stockham (self-sorting) order, so no bit-reversal pass is
// needed.  Twiddles come from a per-plan table generated on the host in FP64
// and rounded once to FP32.
//
// Direction: inverse (unnormalised), X[k] = sum_j x[j] exp(+2 pi i j k / N),
// the direction of the reference's cuFFT call (FIPkernels.cu:679, :741).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace fip {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (+i)
__device__ __forceinline__ float2 cmul_i(float2 a) { return make_float2(-a.y, a.x); }
// a * (-i)
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// LDS padding: one float2 of padding every 32 elements breaks the
// power-of-two strides of the early Stockham stages across banks.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> 5) + 1; }

// ---------------------------------------------------------- radix butterflies
// In-register inverse DFTs of size R: v[s] <- sum_r v[r] exp(+2 pi i r s / R).
template <int R> struct Bfly;

template <> struct Bfly<2> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <> struct Bfly<3> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float s3 = 0.866025403784438646763723170752936183f;   // sqrt(3)/2
        float2 t = cadd(v[1], v[2]);
        float2 u = csub(v[1], v[2]);
        float2 a = v[0];
        v[0] = cadd(a, t);
        float2 m = make_float2(a.x - 0.5f * t.x, a.y - 0.5f * t.y);
        float2 iu = cmul_i(cscale(u, s3));
        v[1] = cadd(m, iu);
        v[2] = csub(m, iu);
    }
};

template <> struct Bfly<4> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = cmul_i(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <> struct Bfly<8> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float r = 0.707106781186547524400844362104849039f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<4>::run(e);
        Bfly<4>::run(o);
        // o[k] *= exp(+2 pi i k / 8)
        float2 o1 = make_float2(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y));
        float2 o2 = cmul_i(o[2]);
        float2 o3 = make_float2(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// Compile-time sine and cosine (Taylor series in FP64; |x| <= pi) for the
// constant rotations inside composite butterflies.
constexpr double kPiD = 3.14159265358979323846264338327950288;
constexpr double ct_sin(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 24; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 24; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}

// a * exp(+2 pi i K / R) with the angle reduced to (-pi, pi]; quarter turns
// are exact sign swaps, odd eighth turns take two multiplies.
template <int R, int K>
__device__ __forceinline__ float2 rot(float2 a)
{
    constexpr int k = ((K % R) + R) % R;
    if constexpr (k == 0) {
        return a;
    } else if constexpr (4 * k == R) {
        return cmul_i(a);
    } else if constexpr (2 * k == R) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (4 * k == 3 * R) {
        return cmul_mi(a);
    } else {
        constexpr int kr = (2 * k > R) ? k - R : k;
        constexpr double ang = 2.0 * kPiD * (double)kr / (double)R;
        if constexpr ((8 * k) % R == 0) {
            constexpr float r = 0.707106781186547524400844362104849039f;
            constexpr float cs = ct_cos(ang) > 0.0 ? 1.0f : -1.0f;
            constexpr float ss = ct_sin(ang) > 0.0 ? 1.0f : -1.0f;
            return make_float2(r * (cs * a.x - ss * a.y), r * (ss * a.x + cs * a.y));
        } else {
            constexpr float c = (float)ct_cos(ang), s = (float)ct_sin(ang);
            return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
        }
    }
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Composite radix R = A * B in registers (Cooley-Tukey): input r = B r1 + r2,
// output s = s1 + A s2; A-point DFTs over r1, rotation by
// exp(+2 pi i r2 s1 / R), then B-point DFTs over r2.
template <int A, int B>
struct BflyAB {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr int R = A * B;
        float2 x[B][A];
        static_for<B>([&](auto r2) {
            static_for<A>([&](auto r1) { x[r2][r1] = v[B * r1 + r2]; });
            Bfly<A>::run(x[r2]);
            static_for<A>([&](auto s1) { x[r2][s1] = rot<R, decltype(r2)::value * decltype(s1)::value>(x[r2][s1]); });
        });
        static_for<A>([&](auto s1) {
            float2 y[B];
            static_for<B>([&](auto r2) { y[r2] = x[r2][s1]; });
            Bfly<B>::run(y);
            static_for<B>([&](auto s2) { v[s1 + A * s2] = y[s2]; });
        });
    }
};
template <> struct Bfly<6> : BflyAB<2, 3> {};
template <> struct Bfly<12> : BflyAB<4, 3> {};
template <> struct Bfly<16> : BflyAB<4, 4> {};
template <> struct Bfly<24> : BflyAB<8, 3> {};

// ----------------------------------------------------------------- FFT plans
// Radix lists for N = 3 * 2^k.  The host twiddle generator walks the same
// list (make_stage_twiddles), so the two stay consistent by construction.  From
// N = 1536 on, a transform gets N/12 threads (12 points per thread): more
// waves per workgroup hide the LDS and barrier latency of each stage.  Each
// stage is one LDS round trip and two barriers, so these lengths take the
// fewest stages that keep a butterfly's registers within occupancy: radix 16
// and 12 in registers (three stages for 768 to 3072, four for 6144 and
// 12288) measured 3-13% faster than radix 8/4/3/2 lists of five or six
// stages.
template <int NT, int... Rs>
struct PlanDef {
    static constexpr int nt = NT;
    static constexpr int nstages = sizeof...(Rs);
    static constexpr int radix[sizeof...(Rs)] = {Rs...};
};
template <int N> struct FftPlan;
#define FIP_PLAN(N_, NT_, ...)                                 \
    template <> struct FftPlan<N_> : PlanDef<NT_, __VA_ARGS__> { \
        static constexpr int n = N_;                           \
    }
FIP_PLAN(48, 64, 8, 2, 3);
FIP_PLAN(96, 64, 8, 4, 3);
FIP_PLAN(192, 64, 8, 8, 3);
FIP_PLAN(384, 64, 8, 8, 2, 3);
FIP_PLAN(768, 64, 16, 12, 4);
FIP_PLAN(1536, 128, 16, 8, 12);
FIP_PLAN(3072, 256, 16, 16, 12);
FIP_PLAN(6144, 512, 16, 16, 12, 2);
FIP_PLAN(12288, 1024, 16, 16, 12, 4);
#undef FIP_PLAN

// Padded LDS offsets of the R elements of a butterfly relative to the padded
// index of its first element, when they are compile-time constants (so the
// LDS instructions carry them as immediates):
//  - loads j + r*NB with NB % 32 == 0: r * (NB + NB/32);
//  - stores base + r*Ns, base = (j/Ns)*Ns*R + (j%Ns): if Ns*R divides 32,
//    base % 32 + r*Ns < 32 and the offset is r*Ns; if Ns % 32 == 0, or Ns
//    divides 32 and 32 divides Ns*R, base % 32 < Ns and the offset is
//    r*Ns + (r*Ns)/32.
template <int NB> struct LoadOff {
    static constexpr bool ok = (NB % 32) == 0;
    static constexpr __host__ __device__ int at(int r) { return r * (NB + NB / 32); }
};
template <int Ns, int R> struct StoreOff {
    static constexpr bool small = (32 % (Ns * R)) == 0;
    static constexpr bool big = (Ns % 32 == 0) || ((32 % Ns) == 0 && ((Ns * R) % 32) == 0);
    static constexpr bool ok = small || big;
    static constexpr __host__ __device__ int at(int r) { return small ? r * Ns : r * Ns + ((r * Ns) >> 5); }
};

// One Stockham stage of radix R with Ns = product of earlier radices.
// tw points at this stage's twiddles laid out [r-1][k], k < Ns.  Threads
// t >= NT of a larger workgroup take no butterfly but reach both barriers.
template <int N, int NT, int R, int Ns>
__device__ __forceinline__ void fft_stage(float2* __restrict__ buf, const float2* __restrict__ tw, int t)
{
    constexpr int NB = N / R;                    // butterflies in this stage
    constexpr int BPT = (NB + NT - 1) / NT;      // butterflies per thread
    constexpr bool exact = (NB % NT) == 0;
    const bool act = t < NT;
    float2 v[BPT][R];
    // Twiddles do not depend on the data: issue their (L1/L2-resident) loads
    // before the LDS loads and the barrier so their latency is hidden.
    float2 w[BPT][R > 1 ? R - 1 : 1];
    if (Ns > 1) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int j = t + b * NT;
            const int k = (exact || j < NB) ? j % Ns : 0;
#pragma unroll
            for (int r = 1; r < R; ++r) w[b][r - 1] = tw[(r - 1) * Ns + k];
        }
    }
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int j = t + b * NT;
        if (act && (exact || j < NB)) {
            if constexpr (LoadOff<NB>::ok) {
                const float2* p = buf + lds_pad(j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = p[LoadOff<NB>::at(r)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = buf[lds_pad(j + r * NB)];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int j = t + b * NT;
        if (act && (exact || j < NB)) {
            const int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[b][r - 1]);
            }
            Bfly<R>::run(v[b]);
            const int base = (j / Ns) * Ns * R + k;
            if constexpr (StoreOff<Ns, R>::ok) {
                float2* p = buf + lds_pad(base);
#pragma unroll
                for (int r = 0; r < R; ++r) p[StoreOff<Ns, R>::at(r)] = v[b][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) buf[lds_pad(base + r * Ns)] = v[b][r];
            }
        }
    }
    __syncthreads();
}

template <class P, int S, int Ns, int TwOff, bool Done = (S >= P::nstages)>
struct FftRun {
    static __device__ __forceinline__ void run(float2* buf, const float2* tw, int t)
    {
        constexpr int R = P::radix[S];
        fft_stage<P::n, P::nt, R, Ns>(buf, tw + TwOff, t);
        constexpr int next_off = TwOff + (Ns > 1 ? (R - 1) * Ns : 0);
        FftRun<P, S + 1, Ns * R, next_off>::run(buf, tw, t);
    }
};

// Terminates the stage recursion once all radices are consumed.
template <class P, int S, int Ns, int TwOff>
struct FftRun<P, S, Ns, TwOff, true> {
    static __device__ __forceinline__ void run(float2*, const float2*, int) {}
};

// Inverse FFT of the N-point padded LDS array buf.  Caller must have a
// workgroup barrier between filling buf and this call; on return the result
// is in buf (natural order) and visible to the whole workgroup.
template <int N>
__device__ __forceinline__ void fft_inverse(float2* buf, const float2* __restrict__ tw, int t)
{
    FftRun<FftPlan<N>, 0, 1, 0>::run(buf, tw, t);
}

// Number of twiddle entries the plan of length N uses (N - 1).
__host__ __device__ constexpr int fft_twiddle_count(int n) { return n - 1; }

}  // namespace fip

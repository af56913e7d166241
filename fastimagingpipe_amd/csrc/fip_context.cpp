// Device context, plan tables, work buffers and the launch sequence.
#include "fip_context.hpp"

#include "fip.h"
#include "fip_plan.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>

namespace fip {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

const char* last_error_cstr() { return g_last_error.c_str(); }

int hip_error(hipError_t e, const char* what)
{
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? FIP_ERR_ALLOC : FIP_ERR_HIP;
    return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

hipError_t DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap && p) return hipSuccess;
    release();
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = bytes;
    return hipSuccess;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

hipError_t PinnedBuf::ensure(size_t bytes)
{
    if (bytes <= cap && p) return hipSuccess;
    release();
    if (bytes == 0) bytes = 256;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = bytes;
    return hipSuccess;
}

void PinnedBuf::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------- work

// Bytes of each work buffer for S snapshots, in the order rec, bkslot,
// sorted, hist, bstart, meta, geom, kt, pre, img, work, pmax.
void work_sizes(int S, size_t stride, const PlanTables& p, size_t out[kWorkBufs])
{
    const size_t s = (size_t)S;
    const size_t nvis = s * stride;
    const size_t sizes[kWorkBufs] = {nvis * sizeof(VisRec),                          // rec
                                     nvis * sizeof(int2),                            // bkslot
                                     nvis * sizeof(VisRec),                          // sorted
                                     s * p.G * sizeof(int),                          // hist
                                     s * (p.G + 1) * sizeof(int),                    // bstart
                                     s * sizeof(SnapMeta),                           // meta
                                     s * sizeof(SnapGeom),                           // geom
                                     s * (size_t)p.I * (p.M + 1) * sizeof(float2),   // kt
                                     s * (size_t)p.I * p.I * sizeof(float),          // pre
                                     s * (size_t)p.I * p.I * sizeof(float),          // img
                                     256,                                            // work
                                     pmax_words(S) * sizeof(unsigned)};              // pmax
    std::copy(sizes, sizes + kWorkBufs, out);
}

size_t Work::bytes(int S, size_t stride, const PlanTables& p) const
{
    size_t sz[kWorkBufs];
    work_sizes(S, stride, p, sz);
    size_t total = 0;
    for (size_t b : sz) total += b;
    return total;
}

int Work::ensure(int S, size_t stride, const PlanTables& p)
{
    DevBuf* const all[kWorkBufs] = {&rec, &bkslot, &sorted, &hist, &bstart, &meta,
                                    &geom, &kt, &pre, &img, &work, &pmax};
    size_t want[kWorkBufs];
    work_sizes(S, stride, p, want);
    for (int i = 0; i < kWorkBufs; ++i) {
        void* const was = all[i]->p;
        const hipError_t e = all[i]->ensure(want[i]);
        if (all[i]->p != was) ++gen;
        if (e != hipSuccess) return hip_error(e, "work buffer allocation");
    }
    return FIP_OK;
}

WorkView Work::view(size_t stride) const
{
    WorkView v;
    v.rec = rec.as<VisRec>();
    v.bkslot = bkslot.as<int2>();
    v.sorted = sorted.as<VisRec>();
    v.hist = hist.as<int>();
    v.bstart = bstart.as<int>();
    v.meta = meta.as<SnapMeta>();
    v.geom = geom.as<SnapGeom>();
    v.kt = kt.as<float2>();
    v.pre = pre.as<float>();
    v.img = img.as<float>();
    v.work = work.as<unsigned>();
    v.pmax = pmax.as<unsigned>();
    v.stride = stride;
    return v;
}

WorkView snapshot_view(const WorkView& w, int s, const PlanTables& p)
{
    const size_t S = (size_t)s;
    WorkView v = w;
    v.rec = w.rec + S * w.stride;
    v.bkslot = w.bkslot + S * w.stride;
    v.sorted = w.sorted + S * w.stride;
    v.hist = w.hist + S * p.G;
    v.bstart = w.bstart + S * (p.G + 1);
    v.meta = w.meta + S;
    v.geom = w.geom + S;
    v.kt = w.kt + S * (size_t)p.I * (p.M + 1);
    v.pre = w.pre + S * (size_t)p.I * p.I;
    v.img = w.img + S * (size_t)p.I * p.I;
    v.pmax = w.pmax + pmax_index(s, 0, 0);
    return v;
}

void Work::release()
{
    for (DevBuf* b : {&rec, &bkslot, &sorted, &hist, &bstart, &meta, &geom, &kt, &pre, &img, &work, &pmax})
        b->release();
    ++gen;
}

// ------------------------------------------------------------- context

int Context::get(Context** out)
{
    static std::mutex init_mu;
    static Context* ctx = nullptr;   // intentionally never destroyed
    std::lock_guard<std::mutex> lk(init_mu);
    if (!ctx) {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            return set_error(FIP_ERR_NODEVICE, "no GPU visible to HIP");
        }
        auto* c = new Context();
        FIP_HIP(hipGetDevice(&c->device));
        FIP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        FIP_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        for (hipEvent_t& e : c->in_ready) FIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        FIP_HIP(hipEventCreateWithFlags(&c->work_done_, hipEventDisableTiming));
        for (auto& pair : c->ev_) {
            FIP_HIP(hipEventCreate(&pair[0]));
            FIP_HIP(hipEventCreate(&pair[1]));
        }
        ctx = c;
    }
    *out = ctx;
    return FIP_OK;
}

void plan_host_tables(size_t image, PlanHostTables& t)
{
    const int G = (int)grid_size_for(image);
    const int M = G / 2;
    t.twg.assign(G, make_float2(0.0f, 0.0f));
    t.twm.assign(M, make_float2(0.0f, 0.0f));
    t.wpost.resize(M);
    make_stage_twiddles(G, t.twg.data());
    make_stage_twiddles(M, t.twm.data());
    for (int k = 0; k < M; ++k) {
        const double a = 2.0 * M_PI * (double)k / (double)G;
        t.wpost[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    t.corr = corr_table(image, (size_t)G);
}

int Context::plan(size_t image, Plan** out)
{
    auto it = plans_.find(image);
    if (it != plans_.end()) {
        *out = it->second.get();
        return FIP_OK;
    }
    if (!image_size_supported(image)) return set_error(FIP_ERR_SIZE, "unsupported image size");
    auto P = std::make_unique<Plan>();
    const int I = (int)image;
    const int G = (int)grid_size_for(image);
    const int M = G / 2;
    P->t.I = I;
    P->t.G = G;
    P->t.M = M;
    PlanHostTables h;
    plan_host_tables(image, h);
    FIP_HIP(P->tw_g.ensure(h.twg.size() * sizeof(float2)));
    FIP_HIP(P->tw_m.ensure(h.twm.size() * sizeof(float2)));
    FIP_HIP(P->wpost.ensure(h.wpost.size() * sizeof(float2)));
    FIP_HIP(P->corr.ensure(h.corr.size() * sizeof(float)));
    FIP_HIP(hipMemcpy(P->tw_g.p, h.twg.data(), h.twg.size() * sizeof(float2), hipMemcpyHostToDevice));
    FIP_HIP(hipMemcpy(P->tw_m.p, h.twm.data(), h.twm.size() * sizeof(float2), hipMemcpyHostToDevice));
    FIP_HIP(hipMemcpy(P->wpost.p, h.wpost.data(), h.wpost.size() * sizeof(float2), hipMemcpyHostToDevice));
    FIP_HIP(hipMemcpy(P->corr.p, h.corr.data(), h.corr.size() * sizeof(float), hipMemcpyHostToDevice));
    P->t.tw_g = P->tw_g.as<float2>();
    P->t.tw_m = P->tw_m.as<float2>();
    P->t.wpost = P->wpost.as<float2>();
    P->t.corr = P->corr.as<float>();
    *out = P.get();
    plans_[image] = std::move(P);
    return FIP_OK;
}

Tuning default_tuning(size_t image)
{
    Tuning t;
    // Gridding records per thread in registers: 2, the autotuner's choice at
    // 4096^2 and 2048^2 on MI355X (each thread's records stay resident up
    // to 1024 per column; heavier columns take further passes).
    t.stage1_records = 2;
    // Rows per C2R workgroup: four rows of up to 768 points, two of 1536 and
    // 3072 (each pair of rows read with one 16-byte load per K entry: 6%
    // faster than four rows at 2048), one of 6144 (two rows of it hold
    // 100 KB of LDS, one workgroup per CU).
    t.stage2_rows = image <= 1024 ? 4 : (image <= 4096 ? 2 : 1);
    t.splat_tile = 64;
    t.persistent = 0;
    return t;
}

Tuning Context::tuning(size_t image) const
{
    auto it = tunings_.find(image);
    if (it != tunings_.end()) return it->second;
    return default_tuning(image);
}

int Context::order_work(hipStream_t st)
{
    if (work_stream_ && work_stream_ != st) FIP_HIP(hipStreamWaitEvent(st, work_done_, 0));
    return FIP_OK;
}

int Context::ensure_win012()
{
    if (win012_ && win.p) return FIP_OK;
    static const int w[3] = {0, 1, 2};
    FIP_HIP(win.ensure(sizeof(w)));
    FIP_HIP(hipMemcpy(win.p, w, sizeof(w), hipMemcpyHostToDevice));
    win012_ = true;
    return FIP_OK;
}

int Context::mark_work(hipStream_t st)
{
    FIP_HIP(hipEventRecord(work_done_, st));
    work_stream_ = st;
    return FIP_OK;
}

int Context::begin_stage(int st, hipStream_t s)
{
    if (!profiling) return FIP_OK;
    FIP_HIP(hipEventRecord(ev_[st][0], s));
    return FIP_OK;
}

int Context::end_stage(int st, hipStream_t s)
{
    if (!profiling) return FIP_OK;
    FIP_HIP(hipEventRecord(ev_[st][1], s));
    ev_used_[st] = true;
    return FIP_OK;
}

int Context::collect_stages()
{
    if (!profiling) return FIP_OK;
    for (int st = 0; st < kStCount; ++st) {
        if (!ev_used_[st]) continue;
        FIP_HIP(hipEventSynchronize(ev_[st][1]));
        float ms = 0.0f;
        FIP_HIP(hipEventElapsedTime(&ms, ev_[st][0], ev_[st][1]));
        stage_ms[st] += ms;
        stage_n[st] += 1;
        ev_used_[st] = false;
    }
    return FIP_OK;
}

// -------------------------------------------------------- launch sequence

// Stage launch with optional profiling events (Cp may be null).
#define FIP_STAGE(Cp, st, s, call)                                 \
    do {                                                           \
        if (Cp) {                                                  \
            int rc_ = (Cp)->begin_stage((st), (s));                \
            if (rc_) return rc_;                                   \
        }                                                          \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return hip_error(e_, stage_name(st)); \
        if (Cp) {                                                  \
            int rc_ = (Cp)->end_stage((st), (s));                  \
            if (rc_) return rc_;                                   \
        }                                                          \
    } while (0)

int imaging_sequence(Context* C, const PlanTables& p, const WorkView& w, const InputView& in, int S, int nmax,
                     float cell, const int* norm_dev, int norm_const, const Tuning& tu, hipStream_t st,
                     int ndefer)
{
    // r1r2_scale = cell * g (FIPkernels.cu:584), kept in FP64.
    const double scale = (double)cell * (double)p.G;
    FIP_STAGE(C, kStReset, st, launch_reset(in, w, p, S, cell, norm_dev, norm_const, st));
    FIP_STAGE(C, kStBin, st, launch_bin(in, w, p, S, nmax, scale, st));
    FIP_STAGE(C, kStScan, st, launch_scan(w, p, S, st));
    FIP_STAGE(C, kStScatter, st, launch_scatter(w, p, S, nmax, st));
    FIP_STAGE(C, kStGridFft, st, launch_stage1(w, p, S, tu, st));
    FIP_STAGE(C, kStRowC2R, st, launch_stage2(w, p, S, tu, st));
    const int ns = S - ndefer;
    if (ns > 0) FIP_STAGE(C, kStSplat, st, launch_splat(w, p, ns, tu, st));
    return FIP_OK;
}

int trigger_sequence(Context* C, const PlanTables& p, const WorkView& w, const int* win_dev, int nwin, int tile,
                     float* out_dev, hipStream_t st)
{
    FIP_STAGE(C, kStTrigger, st, launch_trigger(w, p, win_dev, nwin, tile, out_dev, st));
    return FIP_OK;
}

int fused_trigger_sequence(Context* C, const PlanTables& p, const WorkView& w, int tile, float* out_dev,
                           hipStream_t st)
{
    const size_t plane = (size_t)p.I * p.I;
    FIP_STAGE(C, kStSplatTrig, st, launch_splat_trigger(w, p, 2, w.img, w.img + plane, false, st));
    FIP_STAGE(C, kStTrigger, st, launch_trigger_finish(w, p, 0, 1, 2, tile, out_dev, st));
    return FIP_OK;
}

int pipe_sequence(Context* C, const PlanTables& p, const WorkView& w, const InputView& in, int nmax, float cell,
                  const Tuning& tu, const int* win_dev, int tile, float* out_dev, hipStream_t st)
{
    const bool fuse = splat_trigger_fusable(p, tu, tile);
    if (int rc = imaging_sequence(C, p, w, in, 3, nmax, cell, nullptr, nmax, tu, st, fuse ? 1 : 0)) return rc;
    if (fuse) return fused_trigger_sequence(C, p, w, tile, out_dev, st);
    return trigger_sequence(C, p, w, win_dev, 1, tile, out_dev, st);
}

int run_imaging(Context& C, const Plan& P, Work& W, const InputView& in, int S, int nmax, float cell,
                const int* norm_dev, int norm_const, const Tuning& tu, hipStream_t st)
{
    return imaging_sequence(&C, P.t, W.view(in.stride), in, S, nmax, cell, norm_dev, norm_const, tu, st);
}

int run_trigger(Context& C, const Plan& P, Work& W, size_t stride, const int* win_dev, int nwin, int tile,
                float* out_dev, hipStream_t st)
{
    return trigger_sequence(&C, P.t, W.view(stride), win_dev, nwin, tile, out_dev, st);
}

int run_trigger_run(Context& C, const Plan& P, Work& W, size_t stride, int s0, int nwin, int tile, float* out_dev,
                    hipStream_t st)
{
    const WorkView w = W.view(stride);
    FIP_STAGE(&C, kStTrigger, st, launch_trigger_run(w, P.t, s0, nwin, tile, out_dev, st));
    return FIP_OK;
}

int check_shape(size_t num_baselines, size_t image_size, size_t unit_size, float cell_size)
{
    if (num_baselines == 0) return set_error(FIP_ERR_ARG, "num_baselines must be positive");
    if (num_baselines > (size_t)std::numeric_limits<int>::max() / 4)
        return set_error(FIP_ERR_SIZE, "num_baselines too large");
    if (!std::isfinite(cell_size)) return set_error(FIP_ERR_ARG, "cell_size must be finite");
    if (!image_size_supported(image_size))
        return set_error(FIP_ERR_SIZE,
                         "image_size must be a multiple of 32 with grid 1.5*image = 3*2^k (64..8192), got " +
                             std::to_string(image_size));
    if (unit_size == 0 || unit_size > image_size)
        return set_error(FIP_ERR_SIZE, "unit_size must be in [1, image_size]");
    return FIP_OK;
}

}  // namespace fip

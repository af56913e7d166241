// Serving form: streaming imaging + trigger for several sub-array streams.
//
// Each stream owns a HIP stream (default priority), single-snapshot work
// buffers, a device ring of `ring` image slots and a queue of input slots.
// A submitted snapshot is copied to a free input slot, imaged into the next
// ring slot and, from the third snapshot on, triggered against the two
// previous ring slots (the window of FIPkernels.cu:806-851); the tLISI is
// copied to pinned memory and an event recorded behind it; a completion
// thread polls the events of the snapshots in flight and stamps each
// completion time, so no host callback ever stalls a stream.  Admission: device memory for all streams must fit the HBM
// budget at creation; a snapshot is refused when all of its stream's input
// slots are still in flight.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_plan.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace fip;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct InSlot;
}  // namespace

struct fip_server;

namespace {

struct InSlot {
    DevBuf vr, vi, b, v;
    PinnedBuf host;        // staging for the four input arrays
    PinnedBuf result;      // pinned tLISI of this snapshot
    std::atomic<int> busy{0};
    hipEvent_t done = nullptr;   // recorded after the tLISI copy
    double arrival = 0.0;
    long window = -1;      // window index produced by this snapshot, -1 if none
    int stream_id = 0;
    fip_server* srv = nullptr;
};

struct SubStream {
    fip_stream_config cfg{};
    Plan* plan = nullptr;
    Tuning tu{};
    hipStream_t hs = nullptr;
    DevBuf rec, bkslot, sorted, hist, bstart, kt, pre, work;   // one snapshot
    DevBuf img, meta, geom, pmax;                              // ring
    DevBuf win, out;
    std::vector<std::unique_ptr<InSlot>> slots;
    long submitted = 0;
    std::vector<float> latest;
    long latest_window = -1;
    size_t in_bytes = 0;
    int G = 0, M = 0;
};

}  // namespace

struct fip_server {
    std::vector<std::unique_ptr<SubStream>> streams;
    int ring = 8;
    std::mutex mu;                 // guards latencies / latest / counters
    std::vector<double> lat_ms;
    long processed = 0, refused = 0, inflight = 0;
    // completion tracking
    std::mutex pmu;
    std::condition_variable pcv;
    std::vector<InSlot*> pending;  // submitted, event not yet seen complete
    bool stop = false;
    std::thread poller;
};

namespace {

void on_complete(InSlot* sl, double t)
{
    fip_server* srv = sl->srv;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->lat_ms.push_back((t - sl->arrival) * 1e3);
        srv->processed += 1;
        srv->inflight -= 1;
        if (sl->window >= 0) {
            SubStream& ss = *srv->streams[sl->stream_id];
            const size_t un = ss.cfg.image_size / ss.cfg.unit_size;
            if (sl->window > ss.latest_window) {
                std::memcpy(ss.latest.data(), sl->result.p, un * un * sizeof(float));
                ss.latest_window = sl->window;
            }
        }
    }
    sl->busy.store(0);
}

// Completion thread: polls the events of the snapshots in flight (any
// stream, any order) and completes each as soon as its event has fired.
void poll_loop(fip_server* srv)
{
    std::vector<InSlot*> work;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(srv->pmu);
            srv->pcv.wait(lk, [&] { return srv->stop || !srv->pending.empty(); });
            if (srv->stop && srv->pending.empty()) return;
            work = srv->pending;
        }
        bool any = false;
        for (InSlot* sl : work) {
            const hipError_t q = hipEventQuery(sl->done);
            if (q == hipErrorNotReady) continue;
            // hipSuccess, or an error that will not clear: complete the slot either way.
            const double t = now_s();
            {
                std::lock_guard<std::mutex> lk(srv->pmu);
                srv->pending.erase(std::find(srv->pending.begin(), srv->pending.end(), sl));
            }
            on_complete(sl, t);
            any = true;
        }
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
}

size_t esize(int vt) { return vt == FIP_VIS_F32 ? 4 : 2; }

// Device bytes of one stream with q input slots.
size_t stream_bytes(const fip_stream_config& c, int ring, int q)
{
    const size_t I = c.image_size, N = c.num_baselines;
    const size_t G = grid_size_for(I), M = G / 2;
    const size_t un = I / c.unit_size;
    const size_t work = N * sizeof(VisRec) * 2 + N * sizeof(int2) + G * 4 + (G + 1) * 4 + I * (M + 1) * 8 + I * I * 4;
    const size_t ringb = (size_t)ring * (I * I * 4 + sizeof(SnapMeta) + sizeof(SnapGeom)) + pmax_words(ring) * sizeof(unsigned);
    const size_t in = 2 * N * esize(c.vis_type) + N * 8 + 36;
    return work + ringb + un * un * 4 + 256 + (size_t)q * in;
}

}  // namespace

extern "C" {

int fip_server_create(const fip_stream_config* streams, int nstreams, int ring, size_t hbm_budget_bytes,
                      fip_server** out)
{
    try {
        if (!streams || nstreams <= 0 || !out || ring < 3)
            return set_error(FIP_ERR_ARG, "need >= 1 stream, a ring of >= 3 slots and an output pointer");
        *out = nullptr;
        for (int k = 0; k < nstreams; ++k) {
            const fip_stream_config& c = streams[k];
            if (int rc = check_shape(c.num_baselines, c.image_size, c.unit_size, c.cell_size)) return rc;
            if (c.vis_type != FIP_VIS_F32 && c.vis_type != FIP_VIS_BF16 && c.vis_type != FIP_VIS_F16)
                return set_error(FIP_ERR_ARG, "bad vis_type");
        }
        // Admission at creation: every stream must fit with one input slot;
        // the remaining budget is shared out as extra input slots (up to 8).
        size_t base = 0;
        for (int k = 0; k < nstreams; ++k) base += stream_bytes(streams[k], ring, 1);
        if (base > hbm_budget_bytes)
            return set_error(FIP_ERR_BUDGET, "HBM budget " + std::to_string(hbm_budget_bytes) + " below the " +
                                                 std::to_string(base) + " bytes one snapshot per stream needs");
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        auto* srv = new fip_server();
        srv->ring = ring;
        srv->poller = std::thread(poll_loop, srv);
        size_t spare = hbm_budget_bytes - base;
        for (int k = 0; k < nstreams; ++k) {
            auto ss = std::make_unique<SubStream>();
            ss->cfg = streams[k];
            if (int rc = C->plan(ss->cfg.image_size, &ss->plan)) {
                fip_server_destroy(srv);
                return rc;
            }
            ss->tu = C->tuning(ss->cfg.image_size);
            ss->G = ss->plan->t.G;
            ss->M = ss->plan->t.M;
            const size_t I = ss->cfg.image_size, N = ss->cfg.num_baselines;
            const size_t un = I / ss->cfg.unit_size;
            ss->in_bytes = 2 * N * esize(ss->cfg.vis_type) + N * 8 + 36;
            int q = 1;
            while (q < 8 && spare >= ss->in_bytes) {
                spare -= ss->in_bytes;
                ++q;
            }
            SubStream* ssp = ss.get();
            srv->streams.push_back(std::move(ss));
            auto fail = [&](int rc) {
                fip_server_destroy(srv);
                return rc;
            };
            hipError_t e = hipStreamCreateWithFlags(&ssp->hs, hipStreamNonBlocking);
            if (e != hipSuccess) return fail(hip_error(e, "hipStreamCreate"));
            const size_t G = ssp->G, M = ssp->M;
            for (auto pr : {std::make_pair(&ssp->rec, N * sizeof(VisRec)), std::make_pair(&ssp->sorted, N * sizeof(VisRec)),
                            std::make_pair(&ssp->bkslot, N * sizeof(int2)), std::make_pair(&ssp->hist, G * 4),
                            std::make_pair(&ssp->bstart, (G + 1) * 4), std::make_pair(&ssp->kt, I * (M + 1) * 8),
                            std::make_pair(&ssp->pre, I * I * 4), std::make_pair(&ssp->work, (size_t)256),
                            std::make_pair(&ssp->img, (size_t)ring * I * I * 4),
                            std::make_pair(&ssp->meta, (size_t)ring * sizeof(SnapMeta)),
                            std::make_pair(&ssp->geom, (size_t)ring * sizeof(SnapGeom)),
                            std::make_pair(&ssp->pmax, pmax_words(ring) * sizeof(unsigned)),
                            std::make_pair(&ssp->win, (size_t)3 * sizeof(int) * ring),
                            std::make_pair(&ssp->out, un * un * 4 * (size_t)q)}) {
                e = pr.first->ensure(pr.second);
                if (e != hipSuccess) return fail(hip_error(e, "hipMalloc (server)"));
            }
            // Window table: window w (snapshot index >= 2) uses ring slots
            // (w-2, w-1, w) mod ring; store all ring rotations.
            std::vector<int> wt(3 * ring);
            for (int r = 0; r < ring; ++r) {
                wt[3 * r] = (r + ring - 2) % ring;
                wt[3 * r + 1] = (r + ring - 1) % ring;
                wt[3 * r + 2] = r;
            }
            e = hipMemcpy(ssp->win.p, wt.data(), wt.size() * sizeof(int), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(hip_error(e, "hipMemcpy (server)"));
            for (int s = 0; s < q; ++s) {
                auto sl = std::make_unique<InSlot>();
                sl->srv = srv;
                sl->stream_id = k;
                const size_t vb = N * esize(ssp->cfg.vis_type);
                if ((e = sl->vr.ensure(vb)) != hipSuccess || (e = sl->vi.ensure(vb)) != hipSuccess ||
                    (e = sl->b.ensure(N * 8)) != hipSuccess || (e = sl->v.ensure(36)) != hipSuccess ||
                    (e = sl->host.ensure(2 * vb + N * 8 + 36)) != hipSuccess ||
                    (e = sl->result.ensure(un * un * 4)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&sl->done, hipEventDisableTiming)) != hipSuccess)
                    return fail(hip_error(e, "allocation (server slot)"));
                ssp->slots.push_back(std::move(sl));
            }
            ssp->latest.assign(un * un, 0.0f);
        }
        *out = srv;
        return FIP_OK;
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

int fip_server_submit(fip_server* srv, int sid, const void* vis_re, const void* vis_im, const float* bin,
                      const float* vin, double arrival_s)
{
    if (!srv || !vis_re || !vis_im || !bin || !vin) return set_error(FIP_ERR_ARG, "null argument");
    if (sid < 0 || sid >= (int)srv->streams.size()) return set_error(FIP_ERR_ARG, "bad stream id");
    SubStream& ss = *srv->streams[sid];
    InSlot* sl = nullptr;
    for (auto& s : ss.slots) {
        int expect = 0;
        if (s->busy.compare_exchange_strong(expect, 1)) {
            sl = s.get();
            break;
        }
    }
    if (!sl) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->refused += 1;
        return set_error(FIP_ERR_BUDGET, "all input slots of this stream are in flight");
    }
    Context* C = nullptr;
    if (int rc = Context::get(&C)) {
        sl->busy.store(0);
        return rc;
    }
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    const size_t N = ss.cfg.num_baselines, I = ss.cfg.image_size;
    const size_t vb = N * esize(ss.cfg.vis_type);
    const size_t un = I / ss.cfg.unit_size;
    char* h = static_cast<char*>(sl->host.p);
    std::memcpy(h, vis_re, vb);
    std::memcpy(h + vb, vis_im, vb);
    std::memcpy(h + 2 * vb, bin, N * 8);
    std::memcpy(h + 2 * vb + N * 8, vin, 36);
    sl->arrival = arrival_s >= 0.0 ? arrival_s : now_s();
    const long idx = ss.submitted;
    const int slot = (int)(idx % srv->ring);
    sl->window = idx >= 2 ? idx - 2 : -1;
    auto undo = [&](int rc) {
        sl->busy.store(0);
        return rc;
    };
    hipStream_t st = ss.hs;
#define SRV_HIP(x)                                              \
    do {                                                        \
        hipError_t e_ = (x);                                    \
        if (e_ != hipSuccess) return undo(hip_error(e_, #x));   \
    } while (0)
    SRV_HIP(hipMemcpyAsync(sl->vr.p, h, vb, hipMemcpyHostToDevice, st));
    SRV_HIP(hipMemcpyAsync(sl->vi.p, h + vb, vb, hipMemcpyHostToDevice, st));
    SRV_HIP(hipMemcpyAsync(sl->b.p, h + 2 * vb, N * 8, hipMemcpyHostToDevice, st));
    SRV_HIP(hipMemcpyAsync(sl->v.p, h + 2 * vb + N * 8, 36, hipMemcpyHostToDevice, st));
    WorkView w;
    w.rec = ss.rec.as<VisRec>();
    w.bkslot = ss.bkslot.as<int2>();
    w.sorted = ss.sorted.as<VisRec>();
    w.hist = ss.hist.as<int>();
    w.bstart = ss.bstart.as<int>();
    w.meta = ss.meta.as<SnapMeta>() + slot;
    w.geom = ss.geom.as<SnapGeom>() + slot;
    w.kt = ss.kt.as<float2>();
    w.pre = ss.pre.as<float>();
    w.img = ss.img.as<float>() + (size_t)slot * I * I;
    w.work = ss.work.as<unsigned>();
    w.pmax = ss.pmax.as<unsigned>() + pmax_index(slot, 0, 0);
    w.stride = N;
    InputView in{sl->vr.p, sl->vi.p, sl->b.as<float>(), sl->v.as<float>(), nullptr, N, (int)N, ss.cfg.vis_type};
    const PlanTables& p = ss.plan->t;
    const double scale = (double)ss.cfg.cell_size * (double)p.G;
    SRV_HIP(launch_reset(in, w, p, 1, ss.cfg.cell_size, nullptr, (int)N, st));
    SRV_HIP(launch_bin(in, w, p, 1, (int)N, scale, st));
    SRV_HIP(launch_scan(w, p, 1, st));
    SRV_HIP(launch_scatter(w, p, 1, (int)N, st));
    SRV_HIP(launch_stage1(w, p, 1, ss.tu, st));
    SRV_HIP(launch_stage2(w, p, 1, ss.tu, st));
    SRV_HIP(launch_splat(w, p, 1, ss.tu, st));
    if (sl->window >= 0) {
        WorkView wr = w;
        wr.meta = ss.meta.as<SnapMeta>();
        wr.pmax = ss.pmax.as<unsigned>();
        wr.img = ss.img.as<float>();
        float* dout = ss.out.as<float>();
        SRV_HIP(launch_trigger(wr, p, ss.win.as<int>() + 3 * slot, 1, (int)ss.cfg.unit_size, dout, st));
        SRV_HIP(hipMemcpyAsync(sl->result.p, dout, un * un * 4, hipMemcpyDeviceToHost, st));
    }
    {
        std::lock_guard<std::mutex> lk2(srv->mu);
        srv->inflight += 1;
    }
    hipError_t e = hipEventRecord(sl->done, st);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk2(srv->mu);
        srv->inflight -= 1;
        return undo(hip_error(e, "hipEventRecord"));
    }
    {
        std::lock_guard<std::mutex> lk2(srv->pmu);
        srv->pending.push_back(sl);
    }
    srv->pcv.notify_one();
#undef SRV_HIP
    ss.submitted += 1;
    return FIP_OK;
}

int fip_server_drain(fip_server* srv)
{
    if (!srv) return set_error(FIP_ERR_ARG, "null server");
    for (auto& ss : srv->streams) FIP_HIP(hipStreamSynchronize(ss->hs));
    // Every event has fired; wait for the completion thread to record them.
    for (int spin = 0; spin < 200000; ++spin) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            if (srv->inflight <= 0) return FIP_OK;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
    return set_error(FIP_ERR_HIP, "completions did not arrive");
}

int fip_server_stats(fip_server* srv, double* p50_ms, double* p99_ms, double* mean_ms, long* processed, long* refused)
{
    if (!srv) return set_error(FIP_ERR_ARG, "null server");
    std::vector<double> v;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        v = srv->lat_ms;
        if (processed) *processed = srv->processed;
        if (refused) *refused = srv->refused;
    }
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        std::vector<double> s = v;
        std::sort(s.begin(), s.end());
        const double pos = q * (double)(s.size() - 1);
        const size_t lo = (size_t)std::floor(pos), hi = std::min(s.size() - 1, lo + 1);
        return s[lo] + (pos - (double)lo) * (s[hi] - s[lo]);
    };
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (mean_ms) {
        double m = 0.0;
        for (double x : v) m += x;
        *mean_ms = v.empty() ? 0.0 : m / (double)v.size();
    }
    return FIP_OK;
}

int fip_server_latest(fip_server* srv, int sid, float* tlisi, long* window)
{
    if (!srv || !tlisi) return set_error(FIP_ERR_ARG, "null argument");
    if (sid < 0 || sid >= (int)srv->streams.size()) return set_error(FIP_ERR_ARG, "bad stream id");
    std::lock_guard<std::mutex> lk(srv->mu);
    SubStream& ss = *srv->streams[sid];
    if (ss.latest_window < 0) return set_error(FIP_ERR_ARG, "stream has fewer than 3 images yet");
    std::memcpy(tlisi, ss.latest.data(), ss.latest.size() * sizeof(float));
    if (window) *window = ss.latest_window;
    return FIP_OK;
}

void fip_server_destroy(fip_server* srv)
{
    if (!srv) return;
    (void)fip_server_drain(srv);
    {
        std::lock_guard<std::mutex> lk(srv->pmu);
        srv->stop = true;
    }
    srv->pcv.notify_one();
    if (srv->poller.joinable()) srv->poller.join();
    for (auto& ss : srv->streams) {
        for (DevBuf* b : {&ss->rec, &ss->bkslot, &ss->sorted, &ss->hist, &ss->bstart, &ss->kt, &ss->pre, &ss->work,
                          &ss->img, &ss->meta, &ss->geom, &ss->pmax, &ss->win, &ss->out})
            b->release();
        for (auto& sl : ss->slots) {
            sl->vr.release();
            sl->vi.release();
            sl->b.release();
            sl->v.release();
            sl->host.release();
            sl->result.release();
            if (sl->done) (void)hipEventDestroy(sl->done);
        }
        if (ss->hs) (void)hipStreamDestroy(ss->hs);
    }
    delete srv;
}

}  // extern "C"

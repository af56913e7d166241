// Serving form: streaming imaging + trigger for several sub-array streams.
//
// Each stream owns a HIP stream for its launch sequences and one for its
// input copies (both default priority), work buffers for kServeBatch
// snapshots, a device ring of `ring` image slots and its window table.  A
// submitted snapshot reserves room for its inputs (and its pinned tLISI) in
// an input pool shared by all streams and is copied there through the pool's
// pinned mirror on the copy stream, which the launch sequence that reads it
// waits for.  It is then imaged into its ring slot and, from
// the third snapshot on, triggered against the two previous ring slots (the
// window of FIPkernels.cu:806-851); the tLISI is copied to pinned memory and
// an event recorded behind it.  A completion thread polls the events of the
// snapshots in flight, stamps each completion time and releases the
// reservation, so no host callback ever stalls a stream.
//
// Dynamic batching: a stream keeps at most two launch sequences enqueued
// (the running one and the next, so its GPU queue never runs dry while the
// completion thread relaunches).  A snapshot submitted while two are
// enqueued waits in the stream's queue; when one completes, the completion
// thread launches the queued snapshots (up to kServeBatch, in consecutive
// ring slots) as one multi-snapshot sequence.  An idle stream launches a
// snapshot at once, alone, so the latency at low load is that of one
// snapshot.
//
// Admission by HBM budget: at creation every stream's fixed device state
// (work buffers, ring, window table) plus one snapshot of the largest
// stream's inputs must fit the budget (else FIP_ERR_BUDGET); what remains,
// up to `ring` snapshots per stream, becomes the input pool.  Each arrival is
// admitted only if its reservation fits the pool's free space at that
// moment, so sub-arrays of different sizes share the budget snapshot by
// snapshot; a refused arrival returns FIP_ERR_BUDGET and is counted.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_plan.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <sys/prctl.h>
#include <mutex>
#include <thread>
#include <vector>

using namespace fip;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

size_t esize(int vt) { return vt == FIP_VIS_F32 ? 4 : 2; }

// Snapshots of one stream imaged together when they queue up behind a busy
// stream (work buffers for this many are part of each stream's fixed state).
constexpr int kServeBatch = 3;

// Byte layout of one snapshot's reservation: vis_re | vis_im | bin | vin |
// tLISI (the last only used in the pinned mirror).
struct InLayout {
    size_t vi, b, v, res, bytes;
};

InLayout in_layout(const fip_stream_config& c)
{
    const size_t N = c.num_baselines, un = c.image_size / c.unit_size;
    InLayout L;
    L.vi = al256(N * esize(c.vis_type));
    L.b = L.vi + al256(N * esize(c.vis_type));
    L.v = L.b + al256(N * 2 * sizeof(float));
    L.res = L.v + al256(9 * sizeof(float));
    L.bytes = L.res + al256(un * un * sizeof(float));
    return L;
}

// Inputs of up to kServeBatch queued snapshots gathered for one launch
// sequence, in the InputView layout (stride N): vis_re[B][N] | vis_im[B][N] |
// bin[B][N][2] | vin[B][9], each part 256-byte aligned.
struct BatchLayout {
    size_t vi, b, v, bytes;
};

BatchLayout batch_layout(const fip_stream_config& c)
{
    const size_t N = c.num_baselines, B = kServeBatch, e = esize(c.vis_type);
    BatchLayout L;
    L.vi = al256(B * N * e);
    L.b = L.vi + al256(B * N * e);
    L.v = L.b + al256(B * N * 2 * sizeof(float));
    L.bytes = L.v + al256(B * 9 * sizeof(float));
    return L;
}

// Fixed device bytes of one stream: work buffers and gathered inputs for
// kServeBatch snapshots, the ring of images with their metadata and max
// slots, the window table and the tLISI of kServeBatch windows.
size_t stream_fixed_bytes(const fip_stream_config& c, int ring)
{
    const size_t I = c.image_size, N = c.num_baselines, B = kServeBatch;
    const size_t G = grid_size_for(I), M = G / 2;
    const size_t un = I / c.unit_size;
    const size_t work = N * sizeof(VisRec) * 2 + N * sizeof(int2) + G * 4 + (G + 1) * 4 + I * (M + 1) * 8 + I * I * 4;
    const size_t ringb = (size_t)ring * (I * I * 4 + sizeof(SnapMeta) + sizeof(SnapGeom)) +
                         pmax_words(ring) * sizeof(unsigned) + 3 * sizeof(int) * (size_t)ring;
    return B * work + batch_layout(c).bytes + ringb + B * un * un * 4 + 256;
}

// One launch copies the queued snapshots' inputs from their reservations into
// the batch layout, in 2-byte units (every part is a whole number of them).
struct GatherPart {
    const unsigned short* src;
    unsigned short* dst;
    unsigned n;
};
struct GatherJob {
    GatherPart part[4 * kServeBatch];
    int nparts;
};

__global__ __launch_bounds__(256) void k_gather_inputs(GatherJob job)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < job.nparts; ++k) {
        const GatherPart pt = job.part[k];
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < pt.n; i += stride) pt.dst[i] = pt.src[i];
    }
}

}  // namespace

struct fip_server;

namespace {

struct InSlot {
    size_t off = 0, bytes = 0;   // reservation in the input pool
    hipEvent_t done = nullptr;   // recorded after the tLISI copy
    hipEvent_t copied = nullptr; // recorded after the input copy (copy stream)
    double arrival = 0.0;
    long index = 0;              // position in its stream
    int ring_slot = 0;           // index % ring: its image slot
    long window = -1;            // window index produced by this snapshot, -1 if none
    bool last = false;           // last snapshot of its launch sequence
    int stream_id = 0;
    fip_server* srv = nullptr;
};

struct SubStream {
    fip_stream_config cfg{};
    Plan* plan = nullptr;
    Tuning tu{};
    hipStream_t hs = nullptr;   // launch sequences
    hipStream_t cs = nullptr;   // input copies, so they run under the sequences
    DevBuf rec, bkslot, sorted, hist, bstart, kt, pre, work;   // kServeBatch snapshots
    DevBuf batch_in;                                           // gathered inputs (BatchLayout)
    DevBuf img, meta, geom, pmax;                              // ring
    DevBuf win, out;
    InLayout lay{};
    // Admitted snapshots waiting for a launch sequence, and the sequences
    // enqueued and not yet completed (at most kMaxRunning); both guarded by
    // the server's qmu.
    std::vector<InSlot*> queue;
    int running = 0;
    long sequences = 0;   // launch sequences enqueued (guarded by qmu)
    long submitted = 0;
    std::vector<float> latest;
    long latest_window = -1;
    int G = 0, M = 0;
};

// First-fit allocator over [0, size) with coalescing frees.
class PoolAlloc {
public:
    void reset(size_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool take(size_t bytes, size_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t next = it->first + bytes;
            free_.erase(it);
            if (rest) free_[next] = rest;
            return true;
        }
        return false;
    }
    void give(size_t off, size_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<size_t, size_t> free_;
};

}  // namespace

struct fip_server {
    std::vector<std::unique_ptr<SubStream>> streams;
    int ring = 8;
    std::mutex mu;                 // guards latencies / latest / counters
    std::vector<double> lat_ms;
    long processed = 0, refused = 0, inflight = 0;
    // input pool (device) and its pinned mirror, with the reservations
    DevBuf pool;
    PinnedBuf hpool;
    size_t pool_bytes = 0;
    std::mutex amu;                // guards alloc and idle
    PoolAlloc alloc;
    std::vector<std::unique_ptr<InSlot>> slots;   // every slot object ever made
    std::vector<InSlot*> idle;                    // ready for reuse (event created)
    // completion tracking
    std::mutex pmu;
    std::condition_variable pcv;
    std::vector<InSlot*> pending;  // launched, event not yet seen complete
    bool stop = false;
    std::thread poller;
    std::mutex qmu;                // guards the streams' queues and running counts
    int launch_error = 0;          // first failed launch of the completion thread (under mu)
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
                std::memcpy(ss.latest.data(), static_cast<char*>(srv->hpool.p) + sl->off + ss.lay.res,
                            un * un * sizeof(float));
                ss.latest_window = sl->window;
            }
        }
    }
    std::lock_guard<std::mutex> lk(srv->amu);
    srv->alloc.give(sl->off, sl->bytes);
    srv->idle.push_back(sl);
}

// Completes the slots of a launch sequence that could not be enqueued: the
// stream is synchronised first (copies out of the reservations may be in
// flight), then each reservation is released and the first error kept.
void fail_slots(fip_server* srv, SubStream& ss, const std::vector<InSlot*>& b, int rc)
{
    (void)hipStreamSynchronize(ss.cs);
    (void)hipStreamSynchronize(ss.hs);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->inflight -= (long)b.size();
        if (!srv->launch_error) srv->launch_error = rc;
    }
    std::lock_guard<std::mutex> lk(srv->amu);
    for (InSlot* sl : b) {
        srv->alloc.give(sl->off, sl->bytes);
        srv->idle.push_back(sl);
    }
}

// Enqueues the first queued snapshots of ss (up to kServeBatch, in
// consecutive ring slots) as one launch sequence: their inputs gathered into
// the batch layout (one snapshot is read from its reservation in place), the
// imaging, the windows they complete, the tLISI copies and one event per
// snapshot, which the completion thread then watches.  The caller holds the
// context lock and srv->qmu.
//
// The sequence's images go to ring slots r0 .. r0+k-1, and its first window
// reads the two earlier images in slots r0-2 and r0-1 (mod ring), written by
// earlier sequences.  k is at most ring-2, so no image of the sequence lands
// on those two whatever the order of its splats and triggers, and at most
// ring-r0, so that its slots are consecutive.
int launch_queued(fip_server* srv, SubStream& ss)
{
    if (ss.queue.empty()) return FIP_OK;
    const int ring = srv->ring;
    const int r0 = ss.queue.front()->ring_slot;
    const int k = std::min({(int)ss.queue.size(), kServeBatch, ring - r0, ring - 2});
    std::vector<InSlot*> b(ss.queue.begin(), ss.queue.begin() + k);
    ss.queue.erase(ss.queue.begin(), ss.queue.begin() + k);
    auto fail = [&](int rc) {
        fail_slots(srv, ss, b, rc);
        return rc;
    };
#define SRV_HIP(x)                                              \
    do {                                                        \
        hipError_t e_ = (x);                                    \
        if (e_ != hipSuccess) return fail(hip_error(e_, #x));   \
    } while (0)
    hipStream_t st = ss.hs;
    for (InSlot* sl : b) SRV_HIP(hipStreamWaitEvent(st, sl->copied, 0));
    const size_t N = ss.cfg.num_baselines, I = ss.cfg.image_size, plane = I * I;
    const size_t vb = N * esize(ss.cfg.vis_type);
    const size_t un = I / ss.cfg.unit_size;
    const int T = (int)ss.cfg.unit_size;
    const InLayout& L = ss.lay;
    char* hbase = static_cast<char*>(srv->hpool.p);
    char* dbase = srv->pool.as<char>();
    WorkView w;
    w.rec = ss.rec.as<VisRec>();
    w.bkslot = ss.bkslot.as<int2>();
    w.sorted = ss.sorted.as<VisRec>();
    w.hist = ss.hist.as<int>();
    w.bstart = ss.bstart.as<int>();
    w.meta = ss.meta.as<SnapMeta>() + r0;
    w.geom = ss.geom.as<SnapGeom>() + r0;
    w.kt = ss.kt.as<float2>();
    w.pre = ss.pre.as<float>();
    w.img = ss.img.as<float>() + (size_t)r0 * plane;
    w.work = ss.work.as<unsigned>();
    w.pmax = ss.pmax.as<unsigned>() + pmax_index(r0, 0, 0);
    w.stride = N;
    InputView in;
    if (k == 1) {
        char* d = dbase + b[0]->off;
        in = InputView{d, d + L.vi, reinterpret_cast<const float*>(d + L.b), reinterpret_cast<const float*>(d + L.v),
                       nullptr, N, (int)N, ss.cfg.vis_type};
    } else {
        const BatchLayout BL = batch_layout(ss.cfg);
        char* g = ss.batch_in.as<char>();
        GatherJob job{};
        for (int j = 0; j < k; ++j) {
            const char* d = dbase + b[j]->off;
            const struct {
                const char* src;
                char* dst;
                size_t bytes;
            } parts[4] = {{d, g + j * vb, vb},
                          {d + L.vi, g + BL.vi + j * vb, vb},
                          {d + L.b, g + BL.b + j * N * 8, N * 8},
                          {d + L.v, g + BL.v + j * 36, 36}};
            for (const auto& pt : parts)
                job.part[job.nparts++] = GatherPart{reinterpret_cast<const unsigned short*>(pt.src),
                                                    reinterpret_cast<unsigned short*>(pt.dst),
                                                    (unsigned)(pt.bytes / 2)};
        }
        hipLaunchKernelGGL(k_gather_inputs, dim3(512), dim3(256), 0, st, job);
        SRV_HIP(hipGetLastError());
        in = InputView{g, g + BL.vi, reinterpret_cast<const float*>(g + BL.b), reinterpret_cast<const float*>(g + BL.v),
                       nullptr, N, (int)N, ss.cfg.vis_type};
    }
    const PlanTables& p = ss.plan->t;
    int j0 = 0;   // first snapshot of the sequence that completes a window
    while (j0 < k && b[j0]->window < 0) ++j0;
    // Every snapshot from j0 on completes a window.  With a 32-pixel trigger
    // tile each carries its window's trigger in its own splat, launched one
    // after another (a window's earlier images, from this sequence or an
    // earlier one, are splatted by then), each image still stored for the
    // next two windows; the terms buffer is reused window after window.
    const bool fuse = j0 < k && splat_trigger_fusable(p, ss.tu, T);
    if (int rc = imaging_sequence(nullptr, p, w, in, k, (int)N, ss.cfg.cell_size, nullptr, (int)N, ss.tu, st,
                                  fuse ? k - j0 : 0))
        return fail(rc);
    WorkView wr = w;   // the whole ring
    wr.meta = ss.meta.as<SnapMeta>();
    wr.pmax = ss.pmax.as<unsigned>();
    wr.img = ss.img.as<float>();
    float* dout = ss.out.as<float>();
    if (fuse) {
        for (int j = j0; j < k; ++j) {
            const int sa = (int)((b[j]->index - 2) % ring), sb = (int)((b[j]->index - 1) % ring);
            SRV_HIP(launch_splat_trigger(w, p, j, wr.img + (size_t)sa * plane, wr.img + (size_t)sb * plane, true,
                                         st));
            SRV_HIP(launch_trigger_finish(wr, p, sa, sb, r0 + j, T, dout + (size_t)(j - j0) * un * un, st));
        }
    } else if (j0 < k) {
        // windows (r-2, r-1, r) of ring slots r0+j0 .. r0+k-1, one table row each
        SRV_HIP(launch_trigger(wr, p, ss.win.as<int>() + 3 * (r0 + j0), k - j0, T, dout, st));
    }
    for (int j = j0; j < k; ++j)
        SRV_HIP(hipMemcpyAsync(hbase + b[j]->off + L.res, dout + (size_t)(j - j0) * un * un, un * un * 4,
                               hipMemcpyDeviceToHost, st));
    for (int j = 0; j < k; ++j) {
        b[j]->last = j == k - 1;
        SRV_HIP(hipEventRecord(b[j]->done, st));
    }
#undef SRV_HIP
    ss.running += 1;
    ss.sequences += 1;
    {
        std::lock_guard<std::mutex> lk(srv->pmu);
        srv->pending.insert(srv->pending.end(), b.begin(), b.end());
    }
    srv->pcv.notify_one();
    return FIP_OK;
}

// Launch sequences per stream enqueued at once: the running one and the next.
constexpr int kMaxRunning = 2;

// Launches the queued snapshots of ss while fewer than kMaxRunning
// sequences are enqueued, so the stream's next sequence is always waiting
// behind the running one; what queues up meanwhile goes out together.  The
// caller holds the context lock and srv->qmu.  A failed launch releases the
// rest of the queue unlaunched and returns its error (also kept in
// srv->launch_error).
int launch_ready(fip_server* srv, SubStream& ss)
{
    while (!ss.queue.empty() && ss.running < kMaxRunning) {
        if (int rc = launch_queued(srv, ss)) {
            fail_slots(srv, ss, ss.queue, rc);
            ss.queue.clear();
            return rc;
        }
    }
    return FIP_OK;
}

// After the last snapshot of a stream's launch sequence has completed: the
// queued snapshots of that stream, if any, go out.  If that launch fails, the
// error is kept in srv->launch_error (reported by drain and by later
// submits) and the rest of the queue is released unlaunched.
void relaunch(fip_server* srv, int sid)
{
    SubStream& ss = *srv->streams[sid];
    Context* C = nullptr;
    if (int rc = Context::get(&C)) {
        std::lock_guard<std::mutex> lq(srv->qmu);
        ss.running -= 1;
        fail_slots(srv, ss, ss.queue, rc);
        ss.queue.clear();
        return;
    }
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    std::lock_guard<std::mutex> lq(srv->qmu);
    ss.running -= 1;
    (void)launch_ready(srv, ss);
}

// Completion thread: polls the events of the snapshots in flight (any
// stream, any order) and completes each as soon as its event has fired.
void poll_loop(fip_server* srv)
{
    // The short sleeps between polls are meant literally: with the default
    // timer slack (50 us) each would last ~50 us, and a stream's next launch
    // sequence would wait that long after the previous one completes.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
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
            // read before the slot returns to the idle list
            const bool last = sl->last;
            const int sid = sl->stream_id;
            on_complete(sl, t);
            if (last) relaunch(srv, sid);
            any = true;
        }
        if (!any) std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
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
        // Admission at creation: the fixed state of every stream plus one
        // snapshot of the largest stream must fit; the rest of the budget, up
        // to `ring` snapshots per stream, is the shared input pool.
        size_t base = 0, largest = 0, want = 0;
        for (int k = 0; k < nstreams; ++k) {
            base += stream_fixed_bytes(streams[k], ring);
            const size_t r = in_layout(streams[k]).bytes;
            largest = std::max(largest, r);
            want += (size_t)ring * r;
        }
        if (base + largest > hbm_budget_bytes)
            return set_error(FIP_ERR_BUDGET, "HBM budget " + std::to_string(hbm_budget_bytes) + " below the " +
                                                 std::to_string(base + largest) +
                                                 " bytes the streams and one snapshot need");
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        auto* srv = new fip_server();
        srv->ring = ring;
        srv->pool_bytes = std::min(want, (hbm_budget_bytes - base) / 256 * 256);
        auto fail = [&](int rc) {
            fip_server_destroy(srv);
            return rc;
        };
        hipError_t e = srv->pool.ensure(srv->pool_bytes);
        if (e != hipSuccess) return fail(hip_error(e, "hipMalloc (server input pool)"));
        e = srv->hpool.ensure(srv->pool_bytes);
        if (e != hipSuccess) return fail(hip_error(e, "hipHostMalloc (server input pool)"));
        srv->alloc.reset(srv->pool_bytes);
        srv->poller = std::thread(poll_loop, srv);
        for (int k = 0; k < nstreams; ++k) {
            auto ss = std::make_unique<SubStream>();
            ss->cfg = streams[k];
            ss->lay = in_layout(streams[k]);
            if (int rc = C->plan(ss->cfg.image_size, &ss->plan)) return fail(rc);
            ss->tu = C->tuning(ss->cfg.image_size);
            ss->G = ss->plan->t.G;
            ss->M = ss->plan->t.M;
            const size_t I = ss->cfg.image_size, N = ss->cfg.num_baselines;
            const size_t un = I / ss->cfg.unit_size;
            SubStream* ssp = ss.get();
            srv->streams.push_back(std::move(ss));
            e = hipStreamCreateWithFlags(&ssp->hs, hipStreamNonBlocking);
            if (e != hipSuccess) return fail(hip_error(e, "hipStreamCreate"));
            e = hipStreamCreateWithFlags(&ssp->cs, hipStreamNonBlocking);
            if (e != hipSuccess) return fail(hip_error(e, "hipStreamCreate"));
            const size_t G = ssp->G, M = ssp->M, B = kServeBatch;
            for (auto pr : {std::make_pair(&ssp->rec, B * N * sizeof(VisRec)),
                            std::make_pair(&ssp->sorted, B * N * sizeof(VisRec)),
                            std::make_pair(&ssp->bkslot, B * N * sizeof(int2)), std::make_pair(&ssp->hist, B * G * 4),
                            std::make_pair(&ssp->bstart, B * (G + 1) * 4), std::make_pair(&ssp->kt, B * I * (M + 1) * 8),
                            std::make_pair(&ssp->pre, B * I * I * 4), std::make_pair(&ssp->work, (size_t)256),
                            std::make_pair(&ssp->batch_in, batch_layout(ssp->cfg).bytes),
                            std::make_pair(&ssp->img, (size_t)ring * I * I * 4),
                            std::make_pair(&ssp->meta, (size_t)ring * sizeof(SnapMeta)),
                            std::make_pair(&ssp->geom, (size_t)ring * sizeof(SnapGeom)),
                            std::make_pair(&ssp->pmax, pmax_words(ring) * sizeof(unsigned)),
                            std::make_pair(&ssp->win, (size_t)3 * sizeof(int) * ring),
                            std::make_pair(&ssp->out, B * un * un * 4)}) {
                e = pr.first->ensure(pr.second);
                if (e != hipSuccess) return fail(hip_error(e, "hipMalloc (server)"));
            }
            // Every buffer is written once here, so the first snapshots after
            // creation do not pay the first use of freshly allocated memory.
            for (DevBuf* b : {&ssp->rec, &ssp->bkslot, &ssp->sorted, &ssp->hist, &ssp->bstart, &ssp->kt, &ssp->pre,
                              &ssp->work, &ssp->batch_in, &ssp->img, &ssp->meta, &ssp->geom, &ssp->pmax, &ssp->out}) {
                e = hipMemsetAsync(b->p, 0, b->cap, ssp->hs);
                if (e != hipSuccess) return fail(hip_error(e, "hipMemsetAsync (server)"));
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
            ssp->latest.assign(un * un, 0.0f);
        }
        // The input pool and its pinned mirror are first written here too.
        std::memset(srv->hpool.p, 0, srv->pool_bytes);
        e = hipMemsetAsync(srv->pool.p, 0, srv->pool_bytes, srv->streams[0]->hs);
        if (e != hipSuccess) return fail(hip_error(e, "hipMemsetAsync (server pool)"));
        for (auto& ss : srv->streams) {
            e = hipStreamSynchronize(ss->hs);
            if (e != hipSuccess) return fail(hip_error(e, "hipStreamSynchronize (server)"));
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
    const double t_in = now_s();
    // Admission: this snapshot's reservation must fit the pool's free space now.
    InSlot* sl = nullptr;
    {
        std::lock_guard<std::mutex> lk(srv->amu);
        size_t off = 0;
        if (srv->alloc.take(ss.lay.bytes, &off)) {
            if (srv->idle.empty()) {
                auto fresh = std::make_unique<InSlot>();
                fresh->srv = srv;
                hipError_t e = hipEventCreateWithFlags(&fresh->done, hipEventDisableTiming);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&fresh->copied, hipEventDisableTiming);
                if (e != hipSuccess) {
                    if (fresh->done) (void)hipEventDestroy(fresh->done);
                    srv->alloc.give(off, ss.lay.bytes);
                    return hip_error(e, "hipEventCreate (server)");
                }
                srv->idle.push_back(fresh.get());
                srv->slots.push_back(std::move(fresh));
            }
            sl = srv->idle.back();
            srv->idle.pop_back();
            sl->off = off;
            sl->bytes = ss.lay.bytes;
        }
    }
    if (!sl) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->refused += 1;
        return set_error(FIP_ERR_BUDGET, "input pool of the HBM budget is full: snapshot refused");
    }
    sl->stream_id = sid;
    bool enqueued = false;   // a copy reading the reservation may be in flight
    auto undo = [&](int rc) {
        if (enqueued) (void)hipStreamSynchronize(ss.cs);
        std::lock_guard<std::mutex> lk(srv->amu);
        srv->alloc.give(sl->off, sl->bytes);
        srv->idle.push_back(sl);
        return rc;
    };
    // Host staging into the reservation's pinned mirror needs no lock.
    const size_t N = ss.cfg.num_baselines;
    const size_t vb = N * esize(ss.cfg.vis_type);
    const InLayout& L = ss.lay;
    char* h = static_cast<char*>(srv->hpool.p) + sl->off;
    char* d = srv->pool.as<char>() + sl->off;
    std::memcpy(h, vis_re, vb);
    std::memcpy(h + L.vi, vis_im, vb);
    std::memcpy(h + L.b, bin, N * 8);
    std::memcpy(h + L.v, vin, 36);
    sl->arrival = arrival_s >= 0.0 ? arrival_s : t_in;
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return undo(rc);
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    {
        std::lock_guard<std::mutex> lk2(srv->mu);
        if (srv->launch_error) return undo(srv->launch_error);
    }
    const long idx = ss.submitted;
    sl->index = idx;
    sl->ring_slot = (int)(idx % srv->ring);
    sl->window = idx >= 2 ? idx - 2 : -1;
    // one copy of the contiguous inputs (vis_re .. vin) of the reservation,
    // ordered before any launch sequence that reads it
    enqueued = true;
    {
        hipError_t e = hipMemcpyAsync(d, h, L.res, hipMemcpyHostToDevice, ss.cs);
        if (e == hipSuccess) e = hipEventRecord(sl->copied, ss.cs);
        if (e != hipSuccess) return undo(hip_error(e, "hipMemcpyAsync (server input)"));
    }
    ss.submitted += 1;
    {
        std::lock_guard<std::mutex> lk2(srv->mu);
        srv->inflight += 1;
    }
    std::lock_guard<std::mutex> lq(srv->qmu);
    ss.queue.push_back(sl);
    // launched now, or by the completion thread with the next sequence
    return launch_ready(srv, ss);
}

int fip_server_drain(fip_server* srv)
{
    if (!srv) return set_error(FIP_ERR_ARG, "null server");
    // Queued snapshots are launched by the completion thread when the
    // sequence ahead of them completes, so the streams are synchronised
    // again until the completion thread has recorded every snapshot.
    const double deadline = now_s() + 120.0;
    for (;;) {
        for (auto& ss : srv->streams) FIP_HIP(hipStreamSynchronize(ss->hs));
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            if (srv->inflight <= 0) return srv->launch_error ? set_error(srv->launch_error, "a queued launch failed")
                                                             : FIP_OK;
        }
        if (now_s() > deadline) return set_error(FIP_ERR_HIP, "completions did not arrive");
        std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
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

int fip_server_sequences(fip_server* srv, long* sequences)
{
    if (!srv || !sequences) return set_error(FIP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lq(srv->qmu);
    long n = 0;
    for (auto& ss : srv->streams) n += ss->sequences;
    *sequences = n;
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
                          &ss->batch_in, &ss->img, &ss->meta, &ss->geom, &ss->pmax, &ss->win, &ss->out})
            b->release();
        if (ss->hs) (void)hipStreamDestroy(ss->hs);
        if (ss->cs) (void)hipStreamDestroy(ss->cs);
    }
    for (auto& sl : srv->slots) {
        if (sl->done) (void)hipEventDestroy(sl->done);
        if (sl->copied) (void)hipEventDestroy(sl->copied);
    }
    srv->pool.release();
    srv->hpool.release();
    delete srv;
}

}  // extern "C"

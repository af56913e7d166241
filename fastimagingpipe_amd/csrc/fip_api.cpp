// C ABI of the FIP-TOI path: drop-in pipe, device-resident and precision
// forms, imaging/trigger halves, batching with hipGraph capture, profiling,
// autotuning and the FFT self-check.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_plan.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <random>
#include <vector>

namespace fip {
const char* last_error_cstr();
}

using namespace fip;

namespace {

size_t vis_elem_size(int vt) { return vt == FIP_VIS_F32 ? 4 : 2; }

int check_vis_type(int vt)
{
    if (vt != FIP_VIS_F32 && vt != FIP_VIS_BF16 && vt != FIP_VIS_F16)
        return set_error(FIP_ERR_ARG, "vis_type must be FIP_VIS_F32, FIP_VIS_BF16 or FIP_VIS_F16");
    return FIP_OK;
}

float max3_from_meta(const SnapMeta* m)
{
    float a = 0.0f;
    uint32_t b;
    for (int i = 0; i < 3; ++i) {
        b = m[i].maxd_bits;
        float f;
        std::memcpy(&f, &b, 4);
        a = std::max(a, f);
    }
    return a;
}

// Shared body of fip_pipe / fip_pipe_lowp: host inputs -> host tLISI.
int pipe_host(const void* vis_re, const void* vis_im, int vt, const float* bin, const float* vin, float* result,
              size_t nbl, size_t image, size_t nsnap, float cell, size_t tile)
{
    if (!vis_re || !vis_im || !bin || !vin || !result) return set_error(FIP_ERR_ARG, "null pointer argument");
    if (int rc = check_vis_type(vt)) return rc;
    if (nsnap < 3) return set_error(FIP_ERR_NSNAP, "FIpipe needs at least 3 snapshots");
    if (int rc = check_shape(nbl, image, tile, cell)) return rc;
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return rc;
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    Plan* P = nullptr;
    if (int rc = C->plan(image, &P)) return rc;
    const int S = 3;   // the reference images snapshots 0..2 only (FIPkernels.cu:647)
    const size_t esz = vis_elem_size(vt);
    // One region per snapshot, in pinned staging and on the device:
    // vis_re | vis_im | bin | vin, each part 256-byte aligned.
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t o_vi = al(nbl * esz), o_b = o_vi + al(nbl * esz), o_v = o_b + al(nbl * 2 * sizeof(float));
    const size_t region = o_v + al(9 * sizeof(float));
    const size_t un = image / tile;
    const size_t rbytes = un * un * sizeof(float);
    FIP_HIP(C->staging.ensure(S * region + rbytes + S * sizeof(SnapMeta)));
    char* hs = static_cast<char*>(C->staging.p);
    FIP_HIP(C->in_snap.ensure(S * region));
    char* ds = C->in_snap.as<char>();
    if (int rc = C->work.ensure(S, nbl, P->t)) return rc;
    if (int rc = C->ensure_win012()) return rc;
    FIP_HIP(C->out.ensure(rbytes));
    hipStream_t st = C->stream, cs = C->copy_stream;
    if (int rc = C->order_work(st)) return rc;
    // Snapshot s+1 is copied (host -> pinned staging -> device, on the copy
    // stream) while snapshot s is imaged on the compute stream, as the
    // reference's stream3 does (FIPkernels.cu:648-663, 716-720); the trigger
    // follows the third image.  Within a snapshot the DMA of one part
    // overlaps the host copy of the next.  The staging buffer is free: every
    // return after the first copy below synchronises both streams first.
    char* hr = hs + S * region;
    const Tuning tu = C->tuning(image);
    const WorkView wall = C->work.view(nbl);
    // the third snapshot's splat carries the trigger where it can
    const bool fuse = splat_trigger_fusable(P->t, tu, (int)tile);
    auto enqueue = [&]() -> int {
        for (int s = 0; s < S; ++s) {
            const size_t sv = (size_t)s * nbl;
            const struct {
                const void* src;
                size_t off, bytes;
            } parts[4] = {{static_cast<const char*>(vis_re) + sv * esz, 0, nbl * esz},
                          {static_cast<const char*>(vis_im) + sv * esz, o_vi, nbl * esz},
                          {bin + 2 * sv, o_b, nbl * 2 * sizeof(float)},
                          {vin + 9 * (size_t)s, o_v, 9 * sizeof(float)}};
            char* h = hs + (size_t)s * region;
            char* d = ds + (size_t)s * region;
            for (const auto& pt : parts) {
                std::memcpy(h + pt.off, pt.src, pt.bytes);
                FIP_HIP(hipMemcpyAsync(d + pt.off, h + pt.off, pt.bytes, hipMemcpyHostToDevice, cs));
            }
            FIP_HIP(hipEventRecord(C->in_ready[s], cs));
            FIP_HIP(hipStreamWaitEvent(st, C->in_ready[s], 0));
            InputView in{d, d + o_vi, reinterpret_cast<const float*>(d + o_b), reinterpret_cast<const float*>(d + o_v),
                         nullptr, nbl, (int)nbl, vt};
            if (int rc = imaging_sequence(C, P->t, snapshot_view(wall, s, P->t), in, 1, (int)nbl, cell, nullptr,
                                          (int)nbl, tu, st, fuse && s == S - 1 ? 1 : 0))
                return rc;
        }
        if (fuse) {
            if (int rc = fused_trigger_sequence(C, P->t, wall, (int)tile, C->out.as<float>(), st)) return rc;
        } else if (int rc = run_trigger(*C, *P, C->work, nbl, C->win.as<int>(), 1, (int)tile, C->out.as<float>(),
                                        st)) {
            return rc;
        }
        FIP_HIP(hipMemcpyAsync(hr, C->out.p, rbytes, hipMemcpyDeviceToHost, st));
        FIP_HIP(hipMemcpyAsync(hr + rbytes, C->work.meta.p, S * sizeof(SnapMeta), hipMemcpyDeviceToHost, st));
        return C->mark_work(st);
    };
    const int erc = enqueue();
    // Also on an error: copies out of the staging buffer may be in flight.
    const hipError_t sc = hipStreamSynchronize(cs);
    const hipError_t se = hipStreamSynchronize(st);
    if (erc) return erc;
    if (sc != hipSuccess) return hip_error(sc, "hipStreamSynchronize (copy stream)");
    if (se != hipSuccess) return hip_error(se, "hipStreamSynchronize");
    if (int rc = C->collect_stages()) return rc;
    std::memcpy(result, hr, rbytes);
    SnapMeta meta[3];
    std::memcpy(meta, hr + rbytes, sizeof(meta));
    if (!(max3_from_meta(meta) > 0.0f))
        return set_error(FIP_ERR_DEGENERATE, "all three images are zero: tLISI divides by maxall^2 = 0");
    return FIP_OK;
}

}  // namespace

extern "C" {

const char* fip_version(void) { return "fastimagingpipe_amd 0.1.0 (gfx950)"; }

const char* fip_last_error(void) { return last_error_cstr(); }

const char* fip_status_string(int status)
{
    switch (status) {
    case FIP_OK: return "ok";
    case FIP_ERR_ARG: return "invalid argument";
    case FIP_ERR_NSNAP: return "fewer than 3 snapshots";
    case FIP_ERR_SIZE: return "unsupported size";
    case FIP_ERR_ALLOC: return "allocation failed";
    case FIP_ERR_HIP: return "HIP runtime error";
    case FIP_ERR_NODEVICE: return "no GPU";
    case FIP_ERR_BUDGET: return "HBM budget exceeded";
    case FIP_ERR_IO: return "I/O error";
    case FIP_ERR_DEGENERATE: return "degenerate input (maxall = 0)";
    default: return "unknown status";
    }
}

int fip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

size_t fip_grid_size(size_t image_size)
{
    return image_size_supported(image_size) ? grid_size_for(image_size) : 0;
}

int fip_pipe(const float* vis_re, const float* vis_im, const float* bin, const float* vin, float* result,
             size_t num_baselines, size_t image_size, size_t num_snapshots, float cell_size, size_t unit_size)
{
    try {
        return pipe_host(vis_re, vis_im, FIP_VIS_F32, bin, vin, result, num_baselines, image_size, num_snapshots,
                         cell_size, unit_size);
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

int fip_pipe_lowp(const void* vis_re, const void* vis_im, int vis_type, const float* bin, const float* vin,
                  float* result, size_t num_baselines, size_t image_size, size_t num_snapshots, float cell_size,
                  size_t unit_size)
{
    try {
        return pipe_host(vis_re, vis_im, vis_type, bin, vin, result, num_baselines, image_size, num_snapshots,
                         cell_size, unit_size);
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

int fip_pipe_device(const float* vis_re, const float* vis_im, const float* bin, const float* vin, float* result,
                    size_t num_baselines, size_t image_size, size_t num_snapshots, float cell_size,
                    size_t unit_size, void* stream)
{
    try {
        if (!vis_re || !vis_im || !bin || !vin || !result) return set_error(FIP_ERR_ARG, "null pointer argument");
        if (num_snapshots < 3) return set_error(FIP_ERR_NSNAP, "FIpipe needs at least 3 snapshots");
        if (int rc = check_shape(num_baselines, image_size, unit_size, cell_size)) return rc;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : C->stream;
        if (int rc = C->work.ensure(3, num_baselines, P->t)) return rc;
        if (int rc = C->ensure_win012()) return rc;
        if (int rc = C->order_work(st)) return rc;
        InputView in{vis_re, vis_im, bin, vin, nullptr, num_baselines, (int)num_baselines, FIP_VIS_F32};
        if (int rc = pipe_sequence(C, P->t, C->work.view(num_baselines), in, (int)num_baselines, cell_size,
                                   C->tuning(image_size), C->win.as<int>(), (int)unit_size, result, st))
            return rc;
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

int fip_image(const float* vis_re, const float* vis_im, const float* bin, const float* vin, size_t num_baselines,
              size_t image_size, size_t num_snapshots, float cell_size, float* out_pre, float* out_img,
              float* out_max)
{
    try {
        if (!vis_re || !vis_im || !bin || !vin) return set_error(FIP_ERR_ARG, "null pointer argument");
        if (num_snapshots == 0) return set_error(FIP_ERR_ARG, "num_snapshots must be positive");
        if (int rc = check_shape(num_baselines, image_size, 1, cell_size)) return rc;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        const int S = (int)num_snapshots;
        const size_t nbl = num_baselines;
        const size_t vb = S * nbl * 4, bb = S * nbl * 8, mb = (size_t)S * 36;
        FIP_HIP(C->in_vr.ensure(vb));
        FIP_HIP(C->in_vi.ensure(vb));
        FIP_HIP(C->in_b.ensure(bb));
        FIP_HIP(C->in_v.ensure(mb));
        hipStream_t st = C->stream;
        if (int rc = C->order_work(st)) return rc;
        FIP_HIP(hipStreamSynchronize(st));
        FIP_HIP(hipMemcpy(C->in_vr.p, vis_re, vb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_vi.p, vis_im, vb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_b.p, bin, bb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_v.p, vin, mb, hipMemcpyHostToDevice));
        if (int rc = C->work.ensure(S, nbl, P->t)) return rc;
        InputView in{C->in_vr.p, C->in_vi.p, C->in_b.as<float>(), C->in_v.as<float>(), nullptr, nbl, (int)nbl,
                     FIP_VIS_F32};
        if (int rc = run_imaging(*C, *P, C->work, in, S, (int)nbl, cell_size, nullptr, (int)nbl,
                                 C->tuning(image_size), st))
            return rc;
        FIP_HIP(launch_max_finalize(C->work.view(nbl), S, st));
        if (int rc = C->mark_work(st)) return rc;
        FIP_HIP(hipStreamSynchronize(st));
        if (int rc = C->collect_stages()) return rc;
        const size_t ib = (size_t)S * image_size * image_size * sizeof(float);
        if (out_pre) FIP_HIP(hipMemcpy(out_pre, C->work.pre.p, ib, hipMemcpyDeviceToHost));
        if (out_img) FIP_HIP(hipMemcpy(out_img, C->work.img.p, ib, hipMemcpyDeviceToHost));
        if (out_max) {
            std::vector<SnapMeta> meta(S);
            FIP_HIP(hipMemcpy(meta.data(), C->work.meta.p, S * sizeof(SnapMeta), hipMemcpyDeviceToHost));
            for (int s = 0; s < S; ++s) std::memcpy(&out_max[s], &meta[s].maxd_bits, 4);
        }
        return FIP_OK;
    } catch (const std::bad_alloc&) {
        return set_error(FIP_ERR_ALLOC, "host allocation failed");
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

int fip_trigger(const float* d1, const float* d2, const float* d3, const float maxall[3], size_t image_size,
                size_t unit_size, float* result)
{
    try {
        if (!d1 || !d2 || !d3 || !maxall || !result) return set_error(FIP_ERR_ARG, "null pointer argument");
        if (int rc = check_shape(1, image_size, unit_size, 1.0f)) return rc;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        if (int rc = C->work.ensure(3, 1, P->t)) return rc;
        if (int rc = C->ensure_win012()) return rc;
        if (int rc = C->order_work(C->stream)) return rc;
        FIP_HIP(hipStreamSynchronize(C->stream));
        const size_t ib = image_size * image_size * sizeof(float);
        float* img = C->work.img.as<float>();
        FIP_HIP(hipMemcpy(img, d1, ib, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(img + image_size * image_size, d2, ib, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(img + 2 * image_size * image_size, d3, ib, hipMemcpyHostToDevice));
        SnapMeta meta[3];
        std::memset(meta, 0, sizeof(meta));
        for (int s = 0; s < 3; ++s) std::memcpy(&meta[s].maxd_bits, &maxall[s], 4);
        FIP_HIP(hipMemcpy(C->work.meta.p, meta, sizeof(meta), hipMemcpyHostToDevice));
        // The trigger reads the image maxima from the spread slots: slot 0
        // holds maxall[s], the others zero.
        std::vector<unsigned> slots(pmax_words(3), 0u);
        for (int s = 0; s < 3; ++s) std::memcpy(&slots[pmax_index(s, 1, 0)], &maxall[s], 4);
        FIP_HIP(hipMemcpy(C->work.pmax.p, slots.data(), slots.size() * sizeof(unsigned), hipMemcpyHostToDevice));
        const size_t un = image_size / unit_size;
        FIP_HIP(C->out.ensure(un * un * sizeof(float)));
        if (int rc = run_trigger(*C, *P, C->work, 1, C->win.as<int>(), 1, (int)unit_size, C->out.as<float>(),
                                 C->stream))
            return rc;
        if (int rc = C->mark_work(C->stream)) return rc;
        FIP_HIP(hipStreamSynchronize(C->stream));
        if (int rc = C->collect_stages()) return rc;
        FIP_HIP(hipMemcpy(result, C->out.p, un * un * sizeof(float), hipMemcpyDeviceToHost));
        if (!(std::max(maxall[0], std::max(maxall[1], maxall[2])) > 0.0f))
            return set_error(FIP_ERR_DEGENERATE, "all maxima are zero: tLISI divides by maxall^2 = 0");
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

// ------------------------------------------------------------- batching

namespace {

struct BatchKey {
    const void* vr;
    const void* vi;
    const void* b;
    const void* v;
    void* out;
    void* stream;
    size_t stride, nsnap, norm_n, image, tile;
    int vis_type;
    uint32_t cell_bits;
    int t0, t1, t2, t3;
    int nmax;
    unsigned long long work_gen;   // the work buffers' addresses are baked into the graph
    bool operator==(const BatchKey& o) const { return std::memcmp(this, &o, sizeof(BatchKey)) == 0; }
};

struct BatchGraph {
    bool valid = false;
    BatchKey key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> nvis_cached;
    DevBuf nvis_d, win_d;
    size_t win_for = 0;
};

BatchGraph& batch_graph()
{
    static BatchGraph* g = new BatchGraph();   // intentionally never destroyed
    return *g;
}

}  // namespace

int fip_batch_device(const void* vis_re, const void* vis_im, int vis_type, const float* bin, const float* vin,
                     const int* nvis, size_t stride, size_t nsnap, size_t norm_n, size_t image_size,
                     float cell_size, size_t unit_size, float* d_tlisi, void* stream)
{
    try {
        if (!vis_re || !vis_im || !bin || !vin || !nvis || !d_tlisi)
            return set_error(FIP_ERR_ARG, "null pointer argument");
        if (int rc = check_vis_type(vis_type)) return rc;
        if (nsnap < 3) return set_error(FIP_ERR_NSNAP, "batching needs at least 3 snapshots");
        if (nsnap > 65535) return set_error(FIP_ERR_SIZE, "too many snapshots in one batch");
        int nmax = 0;
        for (size_t s = 0; s < nsnap; ++s) {
            if (nvis[s] < 0 || (size_t)nvis[s] > stride)
                return set_error(FIP_ERR_ARG, "nvis[s] must be in [0, stride]");
            nmax = std::max(nmax, nvis[s]);
        }
        if (int rc = check_shape(std::max<size_t>(stride, 1), image_size, unit_size, cell_size)) return rc;
        if (norm_n == 0) {
            for (size_t s = 0; s < nsnap; ++s)
                if (nvis[s] == 0) return set_error(FIP_ERR_ARG, "norm_n = 0 needs every nvis[s] > 0");
        }
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : C->stream;
        const int S = (int)nsnap;
        if (int rc = C->work.ensure(S, stride, P->t)) return rc;
        BatchGraph& bg = batch_graph();
        // Per-snapshot counts: upload only when they change (stream-ordered
        // before the graph that reads them).
        if (bg.nvis_cached.size() != nsnap || !std::equal(nvis, nvis + nsnap, bg.nvis_cached.begin())) {
            FIP_HIP(hipStreamSynchronize(st));
            FIP_HIP(bg.nvis_d.ensure(nsnap * sizeof(int)));
            FIP_HIP(hipMemcpy(bg.nvis_d.p, nvis, nsnap * sizeof(int), hipMemcpyHostToDevice));
            bg.nvis_cached.assign(nvis, nvis + nsnap);
        }
        if (bg.win_for != nsnap) {
            std::vector<int> w(3 * (nsnap - 2));
            for (size_t k = 0; k + 2 < nsnap; ++k) {
                w[3 * k] = (int)k;
                w[3 * k + 1] = (int)k + 1;
                w[3 * k + 2] = (int)k + 2;
            }
            FIP_HIP(hipStreamSynchronize(st));
            FIP_HIP(bg.win_d.ensure(w.size() * sizeof(int)));
            FIP_HIP(hipMemcpy(bg.win_d.p, w.data(), w.size() * sizeof(int), hipMemcpyHostToDevice));
            bg.win_for = nsnap;
        }
        const Tuning tu = C->tuning(image_size);
        InputView in{vis_re, vis_im, bin, vin, bg.nvis_d.as<int>(), stride, 0, vis_type};
        const int* norm_dev = norm_n == 0 ? bg.nvis_d.as<int>() : nullptr;
        auto enqueue = [&]() -> int {
            if (int rc = run_imaging(*C, *P, C->work, in, S, nmax, cell_size, norm_dev, (int)norm_n, tu, st))
                return rc;
            // windows (k, k+1, k+2): the run form reads each image once per 8 windows
            return run_trigger_run(*C, *P, C->work, stride, 0, S - 2, (int)unit_size, d_tlisi, st);
        };
        if (int rc = C->order_work(st)) return rc;
        if (C->profiling) {
            // Eager launches with per-stage events (no graph).
            if (int rc = enqueue()) return rc;
            if (int rc = C->mark_work(st)) return rc;
            FIP_HIP(hipStreamSynchronize(st));
            return C->collect_stages();
        }
        BatchKey key{};
        std::memset(&key, 0, sizeof(key));
        key.vr = vis_re;
        key.vi = vis_im;
        key.b = bin;
        key.v = vin;
        key.out = d_tlisi;
        key.stream = st;
        key.stride = stride;
        key.nsnap = nsnap;
        key.norm_n = norm_n;
        key.image = image_size;
        key.tile = unit_size;
        key.vis_type = vis_type;
        std::memcpy(&key.cell_bits, &cell_size, 4);
        key.t0 = tu.stage1_records;
        key.t1 = tu.stage2_rows;
        key.t2 = tu.splat_tile;
        key.t3 = tu.persistent;
        key.nmax = nmax;   // sizes the bin/scatter grids baked into the graph
        key.work_gen = C->work.gen;
        if (!bg.valid || !(bg.key == key)) {
            if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
            if (bg.graph) (void)hipGraphDestroy(bg.graph);
            bg.exec = nullptr;
            bg.graph = nullptr;
            bg.valid = false;
            FIP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = enqueue();
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
            bg.graph = g;
            FIP_HIP(hipGraphInstantiate(&bg.exec, bg.graph, nullptr, nullptr, 0));
            bg.key = key;
            bg.valid = true;
        }
        FIP_HIP(hipGraphLaunch(bg.exec, st));
        return C->mark_work(st);
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

int fip_batch_host(const void* vis_re, const void* vis_im, int vis_type, const float* bin, const float* vin,
                   const int* nvis, size_t stride, size_t nsnap, size_t norm_n, size_t image_size,
                   float cell_size, size_t unit_size, float* tlisi)
{
    try {
        if (!vis_re || !vis_im || !bin || !vin || !nvis || !tlisi)
            return set_error(FIP_ERR_ARG, "null pointer argument");
        if (int rc = check_vis_type(vis_type)) return rc;
        if (nsnap < 3) return set_error(FIP_ERR_NSNAP, "batching needs at least 3 snapshots");
        if (int rc = check_shape(std::max<size_t>(stride, 1), image_size, unit_size, cell_size)) return rc;
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        const size_t esz = vis_elem_size(vis_type);
        const size_t vb = nsnap * stride * esz, bb = nsnap * stride * 2 * sizeof(float), mb = nsnap * 9 * sizeof(float);
        const size_t un = image_size / unit_size;
        const size_t ob = (nsnap - 2) * un * un * sizeof(float);
        FIP_HIP(C->in_vr.ensure(vb));
        FIP_HIP(C->in_vi.ensure(vb));
        FIP_HIP(C->in_b.ensure(bb));
        FIP_HIP(C->in_v.ensure(mb));
        FIP_HIP(C->out.ensure(ob));
        if (int rc = C->order_work(C->stream)) return rc;
        FIP_HIP(hipStreamSynchronize(C->stream));
        FIP_HIP(hipMemcpy(C->in_vr.p, vis_re, vb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_vi.p, vis_im, vb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_b.p, bin, bb, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_v.p, vin, mb, hipMemcpyHostToDevice));
        if (int rc = fip_batch_device(C->in_vr.p, C->in_vi.p, vis_type, C->in_b.as<float>(), C->in_v.as<float>(), nvis,
                                      stride, nsnap, norm_n, image_size, cell_size, unit_size, C->out.as<float>(),
                                      nullptr))
            return rc;
        FIP_HIP(hipStreamSynchronize(C->stream));
        FIP_HIP(hipMemcpy(tlisi, C->out.p, ob, hipMemcpyDeviceToHost));
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

// ------------------------------------------------------------- profiling

int fip_set_profiling(int on)
{
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return rc;
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    C->profiling = on != 0;
    return FIP_OK;
}

int fip_stage_count(void) { return kStCount; }

const char* fip_stage_name(int stage) { return stage_name(stage); }

int fip_stage_times(double* ms, long* counts, int n_stages)
{
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return rc;
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    for (int i = 0; i < std::min(n_stages, (int)kStCount); ++i) {
        if (ms) ms[i] = C->stage_ms[i];
        if (counts) counts[i] = C->stage_n[i];
    }
    return FIP_OK;
}

int fip_stage_times_reset(void)
{
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return rc;
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    for (int i = 0; i < kStCount; ++i) {
        C->stage_ms[i] = 0.0;
        C->stage_n[i] = 0;
    }
    return FIP_OK;
}

// --------------------------------------------------------------- tuning

int fip_set_tuning(size_t image_size, const int* choice)
{
    if (!choice) return set_error(FIP_ERR_ARG, "null choice");
    if (!image_size_supported(image_size)) return set_error(FIP_ERR_SIZE, "unsupported image size");
    Context* C = nullptr;
    if (int rc = Context::get(&C)) return rc;
    std::lock_guard<std::recursive_mutex> lk(C->mu);
    Tuning t;
    t.stage1_records = choice[0];
    t.stage2_rows = choice[1];
    t.splat_tile = choice[2];
    t.persistent = choice[3];
    C->set_tuning(image_size, t);
    return FIP_OK;
}

int fip_autotune(size_t image_size, size_t num_baselines, float cell_size, int reps, int* choice)
{
    try {
        if (int rc = check_shape(num_baselines, image_size, 1, cell_size)) return rc;
        if (!(cell_size > 0.0f)) return set_error(FIP_ERR_ARG, "cell_size must be positive");
        reps = std::max(1, reps);
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        Plan* P = nullptr;
        if (int rc = C->plan(image_size, &P)) return rc;
        // Synthetic workload of the caller's shape from the library's own
        // generator: three snapshots of a log-radial array whose longest
        // baseline fills about 54% of the half grid at this cell size
        // (2 R f / c * cell = 0.27), as the plain path images them.
        const size_t nbl = num_baselines;
        int nant = 2;
        while ((size_t)nant * (size_t)(nant - 1) / 2 < nbl) ++nant;
        const size_t nall = (size_t)nant * (size_t)(nant - 1) / 2;
        constexpr int S = 3;
        const double freq = 140e6, c_light = 299792458.0;
        const double radius = 0.135 * c_light / (freq * (double)cell_size);
        std::vector<float> vr(S * nall), vi(S * nall), b(2 * S * nall), v(9 * S);
        if (int rc = fip_synth(nant, radius, S, 1.0, freq, -30.0, -26.7, image_size, cell_size, 20, 0.01, 1.0, -1.8,
                               0.1, 0, nullptr, 0.0, 0.0, 12345ull, vr.data(), vi.data(), b.data(), v.data(), nullptr))
            return rc;
        // First nbl rows of every snapshot, packed at stride nbl.
        std::vector<float> pr(S * nbl), pi(S * nbl), pb(2 * S * nbl);
        for (int s = 0; s < S; ++s) {
            std::copy(vr.begin() + s * nall, vr.begin() + s * nall + nbl, pr.begin() + s * nbl);
            std::copy(vi.begin() + s * nall, vi.begin() + s * nall + nbl, pi.begin() + s * nbl);
            std::copy(b.begin() + 2 * s * nall, b.begin() + 2 * (s * nall + nbl), pb.begin() + 2 * s * nbl);
        }
        if (int rc = C->order_work(C->stream)) return rc;
        FIP_HIP(hipStreamSynchronize(C->stream));
        FIP_HIP(C->in_vr.ensure(S * nbl * 4));
        FIP_HIP(C->in_vi.ensure(S * nbl * 4));
        FIP_HIP(C->in_b.ensure(S * nbl * 8));
        FIP_HIP(C->in_v.ensure(S * 36));
        FIP_HIP(hipMemcpy(C->in_vr.p, pr.data(), S * nbl * 4, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_vi.p, pi.data(), S * nbl * 4, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_b.p, pb.data(), S * nbl * 8, hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(C->in_v.p, v.data(), S * 36, hipMemcpyHostToDevice));
        if (int rc = C->work.ensure(S, nbl, P->t)) return rc;
        InputView in{C->in_vr.p, C->in_vi.p, C->in_b.as<float>(), C->in_v.as<float>(), nullptr, nbl, (int)nbl,
                     FIP_VIS_F32};
        const bool was = C->profiling;
        C->profiling = true;
        double saved_ms[kStCount];
        long saved_n[kStCount];
        std::copy(C->stage_ms, C->stage_ms + kStCount, saved_ms);
        std::copy(C->stage_n, C->stage_n + kStCount, saved_n);
        auto measure = [&](const Tuning& t, double* out) -> int {
            for (int i = 0; i < kStCount; ++i) {
                C->stage_ms[i] = 0.0;
                C->stage_n[i] = 0;
            }
            for (int r = 0; r < reps + 1; ++r) {   // run 0 is a warm-up
                if (int rc = run_imaging(*C, *P, C->work, in, S, (int)nbl, cell_size, nullptr, (int)nbl, t,
                                         C->stream))
                    return rc;
                FIP_HIP(hipStreamSynchronize(C->stream));
                if (int rc = C->collect_stages()) return rc;
                if (r == 0) {
                    for (int i = 0; i < kStCount; ++i) {
                        C->stage_ms[i] = 0.0;
                        C->stage_n[i] = 0;
                    }
                }
            }
            for (int i = 0; i < kStCount; ++i) out[i] = C->stage_ms[i] / std::max(1L, C->stage_n[i]);
            return FIP_OK;
        };
        Tuning best = C->tuning(image_size);
        double best_s1 = 1e300, best_s2 = 1e300, best_sp = 1e300;
        const int records[3] = {2, 4, 8};
        const int rows[3] = {1, 2, 4};
        const int tiles[2] = {32, 64};
        double tm[kStCount];
        for (int ci = 0; ci < 3; ++ci) {
            for (int pers = 0; pers < 2; ++pers) {
                Tuning t = best;
                t.stage1_records = records[ci];
                t.persistent = pers;
                if (int rc = measure(t, tm)) return rc;
                if (tm[kStGridFft] < best_s1) {
                    best_s1 = tm[kStGridFft];
                    best.stage1_records = t.stage1_records;
                    best.persistent = t.persistent;
                }
            }
        }
        for (int r : rows) {
            Tuning t = best;
            t.stage2_rows = r;
            if (int rc = measure(t, tm)) return rc;
            if (tm[kStRowC2R] < best_s2) {
                best_s2 = tm[kStRowC2R];
                best.stage2_rows = r;
            }
        }
        for (int ts : tiles) {
            Tuning t = best;
            t.splat_tile = ts;
            if (int rc = measure(t, tm)) return rc;
            if (tm[kStSplat] < best_sp) {
                best_sp = tm[kStSplat];
                best.splat_tile = ts;
            }
        }
        std::copy(saved_ms, saved_ms + kStCount, C->stage_ms);
        std::copy(saved_n, saved_n + kStCount, C->stage_n);
        C->profiling = was;
        if (int rc = C->mark_work(C->stream)) return rc;
        C->set_tuning(image_size, best);
        if (choice) {
            choice[0] = best.stage1_records;
            choice[1] = best.stage2_rows;
            choice[2] = best.splat_tile;
            choice[3] = best.persistent;
        }
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

// ------------------------------------------------------------ FFT check

int fip_debug_fft(int n, float* data, int count)
{
    try {
        if (!data || count <= 0) return set_error(FIP_ERR_ARG, "bad arguments");
        if (!fft_size_supported(n)) return set_error(FIP_ERR_SIZE, "no compiled FFT plan for this length");
        Context* C = nullptr;
        if (int rc = Context::get(&C)) return rc;
        std::lock_guard<std::recursive_mutex> lk(C->mu);
        std::vector<float2> tw(std::max(1, n - 1));
        make_stage_twiddles(n, tw.data());
        DevBuf dtw, dd;
        FIP_HIP(dtw.ensure(tw.size() * sizeof(float2)));
        const size_t bytes = (size_t)count * n * sizeof(float2);
        FIP_HIP(dd.ensure(bytes));
        FIP_HIP(hipMemcpy(dtw.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
        FIP_HIP(hipMemcpy(dd.p, data, bytes, hipMemcpyHostToDevice));
        FIP_HIP(launch_fft_check(n, dd.as<float2>(), dtw.as<float2>(), count, C->stream));
        FIP_HIP(hipStreamSynchronize(C->stream));
        FIP_HIP(hipMemcpy(data, dd.p, bytes, hipMemcpyDeviceToHost));
        dtw.release();
        dd.release();
        return FIP_OK;
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

void fip_free(void* p) { std::free(p); }

}  // extern "C"

// C++-linkage drop-in with the reference's signature (pipe_interface.cpp:8).
int FIpipe(float* Visreal, float* Visimag, float* Bin, float* Vin, float* result_array, size_t num_baselines,
           size_t image_size, size_t num_snapshots, float cell_size, size_t unit_size)
{
    const int rc = fip_pipe(Visreal, Visimag, Bin, Vin, result_array, num_baselines, image_size, num_snapshots,
                            cell_size, unit_size);
    if (rc != FIP_OK) std::fprintf(stderr, "FIpipe: %s (%s)\n", fip_status_string(rc), fip_last_error());
    return rc;
}

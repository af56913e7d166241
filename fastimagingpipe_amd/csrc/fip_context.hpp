// Device context: stream, per-size plan tables, work buffers and timing.
#pragma once

#include "fip_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace fip {

// Records msg for fip_last_error() and returns code.
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);

#define FIP_HIP(expr)                                   \
    do {                                                \
        hipError_t fip_e_ = (expr);                     \
        if (fip_e_ != hipSuccess) return hip_error(fip_e_, #expr); \
    } while (0)

// A grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
};

struct Plan {
    PlanTables t{};
    DevBuf tw_g, tw_m, wpost, corr;
};

// Host copies of one image size's plan tables (FP64-computed, FP32-stored):
// stage twiddles of the length-G and length-M transforms, the C2R split
// factors and the per-axis grid correction.
struct PlanHostTables {
    std::vector<float2> twg, twm, wpost;
    std::vector<float> corr;
};
void plan_host_tables(size_t image, PlanHostTables& t);

// Launch parameters used until the autotuner or fip_set_tuning picks others.
Tuning default_tuning(size_t image);

// Work buffers in the order rec, bkslot, sorted, hist, bstart, meta, geom,
// kt, pre, img, work, pmax: their byte sizes for S snapshots.
constexpr int kWorkBufs = 12;
void work_sizes(int S, size_t stride, const PlanTables& p, size_t out[kWorkBufs]);

// Work buffers sized for S snapshots of one image size and stride.
struct Work {
    DevBuf rec, bkslot, sorted, hist, bstart, meta, geom, kt, pre, img, work, pmax;
    // Bumped whenever any buffer is reallocated or released: a captured graph
    // that baked in the old addresses must not be replayed after that.
    unsigned long long gen = 0;
    WorkView view(size_t stride) const;
    int ensure(int S, size_t stride, const PlanTables& p);
    size_t bytes(int S, size_t stride, const PlanTables& p) const;
    void release();
};

class Context {
public:
    static int get(Context** out);   // process-wide context on the current device

    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;

    int plan(size_t image, Plan** out);
    Tuning tuning(size_t image) const;
    void set_tuning(size_t image, const Tuning& t) { tunings_[image] = t; }

    // Profiling: event pairs around each stage launch of the current call.
    bool profiling = false;
    double stage_ms[kStCount] = {};
    long stage_n[kStCount] = {};
    int begin_stage(int st, hipStream_t s);
    int end_stage(int st, hipStream_t s);
    int collect_stages();   // call after the stream has drained

    Work work;
    DevBuf in_vr, in_vi, in_b, in_v, out, win, nvis, norm;
    PinnedBuf staging;
    // Host drop-in: per-snapshot input regions on the device, filled on
    // copy_stream; in_ready[s] marks snapshot s's region as copied.
    DevBuf in_snap;
    hipStream_t copy_stream = nullptr;
    hipEvent_t in_ready[3] = {};
    // Preprocessing (projection, collapse and the raw-input chain): device
    // inputs of the host-array entry points, per-block moment partials,
    // per-snapshot projection parameters, and pinned memory for the partials
    // read back and the parameters sent.
    DevBuf pp_uvw, pp_freq, pp_in, pp_part, pp_par;
    PinnedBuf pp_host;

    // Every entry point that enqueues work on `work` from some stream calls
    // order_work(st) before and mark_work(st) after, so uses of the shared
    // buffers from different streams (the context's own and callers') run
    // one after another in call order.
    int order_work(hipStream_t st);
    // C.win holds the window {0, 1, 2} (uploaded once).
    int ensure_win012();
    int mark_work(hipStream_t st);

private:
    std::map<size_t, std::unique_ptr<Plan>> plans_;
    std::map<size_t, Tuning> tunings_;
    hipEvent_t ev_[kStCount][2] = {};
    bool ev_used_[kStCount] = {};
    hipEvent_t work_done_ = nullptr;     // recorded after the last enqueue on `work`
    hipStream_t work_stream_ = nullptr;  // the stream of that enqueue
    bool win012_ = false;
};

// The work buffers of snapshot s of `w` as a one-snapshot view.
WorkView snapshot_view(const WorkView& w, int s, const PlanTables& p);

// Enqueue reset..splat / the trigger on explicit tables and buffers; C (may
// be null) only adds the profiling events.  The splats of the view's last
// ndefer snapshots are left to the caller's fused trigger launches
// (fused_trigger_sequence for FIpipe's third snapshot).
int imaging_sequence(Context* C, const PlanTables& p, const WorkView& w, const InputView& in, int S, int nmax,
                     float cell, const int* norm_dev, int norm_const, const Tuning& tu, hipStream_t st,
                     int ndefer = 0);
int trigger_sequence(Context* C, const PlanTables& p, const WorkView& w, const int* win_dev, int nwin, int tile,
                     float* out_dev, hipStream_t st);
// The fused tail of window (0, 1, 2) of a three-snapshot view whose last
// splat imaging_sequence deferred (splat_trigger_fusable holds).
int fused_trigger_sequence(Context* C, const PlanTables& p, const WorkView& w, int tile, float* out_dev,
                           hipStream_t st);
// FIpipe's three snapshots to the tLISI of window (0, 1, 2) (win_dev holds
// it): the imaging, then the fused tail where splat_trigger_fusable holds,
// else the last splat and the trigger.
int pipe_sequence(Context* C, const PlanTables& p, const WorkView& w, const InputView& in, int nmax, float cell,
                  const Tuning& tu, const int* win_dev, int tile, float* out_dev, hipStream_t st);

// Enqueue reset..splat for S snapshots described by `in` into `w`.
int run_imaging(Context& C, const Plan& P, Work& W, const InputView& in, int S, int nmax, float cell,
                const int* norm_dev, int norm_const, const Tuning& tu, hipStream_t st);
// Enqueue the trigger over windows `win_dev` (3 ints each) into out_dev.
int run_trigger(Context& C, const Plan& P, Work& W, size_t stride, const int* win_dev, int nwin, int tile,
                float* out_dev, hipStream_t st);

// Enqueue the trigger over the consecutive windows (s0+k, s0+k+1, s0+k+2),
// k < nwin.
int run_trigger_run(Context& C, const Plan& P, Work& W, size_t stride, int s0, int nwin, int tile, float* out_dev,
                    hipStream_t st);

// Shape checks shared by every entry point.
int check_shape(size_t num_baselines, size_t image_size, size_t unit_size, float cell_size);

}  // namespace fip

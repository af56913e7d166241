// fip_pipe_device_noalloc: the device-resident FIpipe form for callers that
// cannot allocate device memory (code under stream capture, serving loops
// that forbid hipMalloc).  Every intermediate buffer and every plan table
// lives in one file-scope device array, `scratch`; the call itself allocates
// nothing.
//
// Layout of the array (offsets 256-byte aligned):
//   [tables]  stage twiddles, C2R split factors and grid-correction tables of
//             every image size whose work can fit (64..1024), and the window
//             {0, 1, 2}; written once per device when the device is set up;
//   [work]    the work buffers of one call (rec, bkslot, sorted, hist,
//             bstart, meta, geom, kt, pre, img, work, pmax), laid out per
//             call for its shape.
// Kernels clear what must start at zero (histograms, metadata, max slots)
// in their first launch, so the contents left by the previous call do not
// matter.  Calls on one device share the array: calls that overlap in time
// on one GPU (different streams, or different host threads) race.
#include "fip.h"

#include "fip_context.hpp"
#include "fip_plan.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

namespace fip {
namespace {

// 64 MiB (67.1 MB): the 1024^2 workload with 130816 baselines needs 66.1 MB
// of work buffers (records 22 MB, K tiles 18.9 MB, two image sets 25.2 MB).
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;
// Image sizes whose tables are kept in the array (larger ones cannot fit
// their two image sets).
constexpr size_t kTableImages[] = {64, 128, 256, 512, 1024};
constexpr int kNumTableImages = sizeof(kTableImages) / sizeof(kTableImages[0]);

}  // namespace
}  // namespace fip

// Device memory of the no-allocation entry point; each GPU holds its own copy
// until the process ends.  Host code never uses this name as an address: the
// device address comes from hipGetSymbolAddress (scratch_base below).
static __device__ __align__(256) unsigned char scratch[fip::kScratchBytes];

namespace fip {
namespace {

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct TableSlot {
    size_t twg, twm, wpost, corr;   // byte offsets in the array
};

struct TableLayout {
    TableSlot slot[kNumTableImages];
    size_t win;     // offset of int[3] = {0, 1, 2}
    size_t end;     // first byte after the tables
};

TableLayout table_layout()
{
    TableLayout L{};
    size_t o = 0;
    for (int k = 0; k < kNumTableImages; ++k) {
        const size_t I = kTableImages[k];
        const size_t G = grid_size_for(I), M = G / 2;
        L.slot[k].twg = o;
        o = align_up(o + G * sizeof(float2));
        L.slot[k].twm = o;
        o = align_up(o + M * sizeof(float2));
        L.slot[k].wpost = o;
        o = align_up(o + M * sizeof(float2));
        L.slot[k].corr = o;
        o = align_up(o + (I / 2 + 1) * sizeof(float));
    }
    L.win = o;
    o = align_up(o + 3 * sizeof(int));
    L.end = o;
    return L;
}

struct DeviceScratch {
    unsigned char* base = nullptr;   // device address of this GPU's `scratch`
    bool ready = false;
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// Once per device, with that device current: look up the array's device
// address and write the plan tables into it (synchronous copies, before any
// kernel of this entry point can read them).
int scratch_base(int device, unsigned char** out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(FIP_ERR_NODEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[device];
    if (!d.ready) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return set_error(FIP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        unsigned char* base = static_cast<unsigned char*>(ptr);
        const TableLayout L = table_layout();
        for (int k = 0; k < kNumTableImages; ++k) {
            PlanHostTables h;
            plan_host_tables(kTableImages[k], h);
            FIP_HIP(hipMemcpy(base + L.slot[k].twg, h.twg.data(), h.twg.size() * sizeof(float2),
                              hipMemcpyHostToDevice));
            FIP_HIP(hipMemcpy(base + L.slot[k].twm, h.twm.data(), h.twm.size() * sizeof(float2),
                              hipMemcpyHostToDevice));
            FIP_HIP(hipMemcpy(base + L.slot[k].wpost, h.wpost.data(), h.wpost.size() * sizeof(float2),
                              hipMemcpyHostToDevice));
            FIP_HIP(hipMemcpy(base + L.slot[k].corr, h.corr.data(), h.corr.size() * sizeof(float),
                              hipMemcpyHostToDevice));
        }
        const int win[3] = {0, 1, 2};
        FIP_HIP(hipMemcpy(base + L.win, win, sizeof(win), hipMemcpyHostToDevice));
        d.base = base;
        d.ready = true;
    }
    *out = d.base;
    return FIP_OK;
}

}  // namespace
}  // namespace fip

using namespace fip;

extern "C" int fip_pipe_device_noalloc(const float* vis_re, const float* vis_im, const float* bin, const float* vin,
                                       float* result, size_t num_baselines, size_t image_size, size_t num_snapshots,
                                       float cell_size, size_t unit_size, void* stream)
{
    try {
        if (!vis_re || !vis_im || !bin || !vin || !result) return set_error(FIP_ERR_ARG, "null pointer argument");
        if (num_snapshots < 3) return set_error(FIP_ERR_NSNAP, "FIpipe needs at least 3 snapshots");
        if (int rc = check_shape(num_baselines, image_size, unit_size, cell_size)) return rc;
        int k = -1;
        for (int i = 0; i < kNumTableImages; ++i)
            if (kTableImages[i] == image_size) k = i;
        if (k < 0)
            return set_error(FIP_ERR_SIZE, "fip_pipe_device_noalloc: image_size " + std::to_string(image_size) +
                                               " exceeds its 64 MiB scratch (image sizes 64..1024 only)");
        // Work layout for this shape, after the tables.
        const TableLayout L = table_layout();
        PlanTables p{};
        p.I = (int)image_size;
        p.G = (int)grid_size_for(image_size);
        p.M = p.G / 2;
        size_t sz[kWorkBufs];
        work_sizes(3, num_baselines, p, sz);
        size_t off[kWorkBufs];
        size_t o = L.end;
        for (int i = 0; i < kWorkBufs; ++i) {
            off[i] = o;
            o = align_up(o + sz[i]);
        }
        if (o > kScratchBytes)
            return set_error(FIP_ERR_SIZE, "fip_pipe_device_noalloc: " + std::to_string(num_baselines) +
                                               " baselines at " + std::to_string(image_size) + "^2 need " +
                                               std::to_string(o) + " bytes of scratch, more than its " +
                                               std::to_string(kScratchBytes));
        int dev = 0;
        FIP_HIP(hipGetDevice(&dev));
        unsigned char* base = nullptr;
        if (int rc = scratch_base(dev, &base)) return rc;
        p.tw_g = reinterpret_cast<float2*>(base + L.slot[k].twg);
        p.tw_m = reinterpret_cast<float2*>(base + L.slot[k].twm);
        p.wpost = reinterpret_cast<float2*>(base + L.slot[k].wpost);
        p.corr = reinterpret_cast<float*>(base + L.slot[k].corr);
        WorkView w{};
        w.rec = reinterpret_cast<VisRec*>(base + off[0]);
        w.bkslot = reinterpret_cast<int2*>(base + off[1]);
        w.sorted = reinterpret_cast<VisRec*>(base + off[2]);
        w.hist = reinterpret_cast<int*>(base + off[3]);
        w.bstart = reinterpret_cast<int*>(base + off[4]);
        w.meta = reinterpret_cast<SnapMeta*>(base + off[5]);
        w.geom = reinterpret_cast<SnapGeom*>(base + off[6]);
        w.kt = reinterpret_cast<float2*>(base + off[7]);
        w.pre = reinterpret_cast<float*>(base + off[8]);
        w.img = reinterpret_cast<float*>(base + off[9]);
        w.work = reinterpret_cast<unsigned*>(base + off[10]);
        w.pmax = reinterpret_cast<unsigned*>(base + off[11]);
        w.stride = num_baselines;
        hipStream_t st = static_cast<hipStream_t>(stream);
        InputView in{vis_re, vis_im, bin, vin, nullptr, num_baselines, (int)num_baselines, FIP_VIS_F32};
        return pipe_sequence(nullptr, p, w, in, (int)num_baselines, cell_size, default_tuning(image_size),
                             reinterpret_cast<const int*>(base + L.win), (int)unit_size, result, st);
    } catch (...) {
        return set_error(FIP_ERR_HIP, "unexpected exception");
    }
}

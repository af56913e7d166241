// Internal types and kernel launchers of the FIP-TOI MI355X path.
// Not part of the public C ABI (include/fip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fip {

// ES gridding kernel shape parameter and support (FIPkernels.cu:198, 205).
constexpr double kBeta = 15.3704324328;
constexpr int kSupport = 8;
// K tile height: stage 1 writes 8 consecutive image rows of one uv column as
// one 64-byte run, stage 2 reads the rows of its workgroup from each run
// (measured best among 2, 4, 8 and 16 rows; C2R rows per workgroup must
// divide it).
constexpr int kKTile = 8;

// One binned visibility: integer floor and FP32 fraction of each grid
// coordinate (computed from an FP64 product B * cell * g), and the weighted
// visibility.  24 bytes.
struct VisRec {
    int b1;
    float f1;
    int b2;
    float f2;
    float vr;
    float vi;
};

// Image maxima are gathered in spread slots: per snapshot, kMaxSlots slots
// for the corrected image before re-projection (slot set 0) and kMaxSlots for
// the re-projected image (set 1).  Workgroup g updates slot g % kMaxSlots, so
// about 1/kMaxSlots of a launch's workgroups contend for one address (one
// address for every workgroup serialised the splat's tail: 4096 atomics per
// image cost ~40 us); consumers take the max of the slots.  Contention is per
// address, not per cache line (slots 128 bytes apart measured the same for
// the writers), so the slots are packed and a wave reads them in two lines.
constexpr int kMaxSlots = 64;
// The slot layout is defined once, in the kernels' translation unit: host
// code sizes and fills slot arrays only through these two functions, so an
// allocation can never disagree with the kernels' indexing.
size_t pmax_words(int S);                       // unsigned words for S snapshots
size_t pmax_index(int s, int set, int slot);    // word of one slot

// Per-snapshot scalars produced and consumed on the device.
struct SnapMeta {
    int ngrid;            // visibilities with at least one grid tap
    int rmax;             // largest |column tap|: stage 1 covers c = 0..rmax
    int r1max;            // largest |row tap|: rows of the gridded band
    unsigned maxv_bits;   // max(|re|,|im|) of weighted visibilities (float bits)
    unsigned maxp_bits;   // max of the corrected image before re-projection (k_max_finalize)
    unsigned maxd_bits;   // max of the re-projected image, maxall entry (k_trigger, k_max_finalize)
    int norm_n;           // N used by the splat's 1/N (num_baselines)
    int nvis;             // visibilities present in this snapshot
};

// Per-snapshot re-projection geometry (FP64, from the V matrix).
struct SnapGeom {
    double m00, m01, m10, m11;   // affine: (a,b) = source - h  ->  (q1,q2) = target - h
    double i00, i01, i10, i11;   // inverse of the affine part
    double im00, im10;           // 1/m00, 1/m10 (0 where |m| <= 1e-12)
    double sx, sy;               // SIN shift per unit n' (pixels): -xi/cell, eta/cell
    double delta;                // cell (rad)
    double curv;                 // bound on the SIN map's curvature (1/px), <0: unbounded
    double kappa;                // bound on |grad| of the SIN shift over the image
    int sin_on;                  // xi^2 + eta^2 != 0
    int series;                  // every direction cosine has r^2 <= 0.01 and the splat's quadratic
                                 // model of the series along a group is within 5e-8 px: n' by its series
    int ccap;                    // upper bound on the sources whose splat reaches one pixel
    int pad;
};

enum VisType { kVisF32 = 0, kVisBF16 = 1, kVisF16 = 2 };

// Device-side view of the inputs of a batch of snapshots.
struct InputView {
    const void* vis_re;     // [s*stride + n] of VisType
    const void* vis_im;
    const float* bin;       // [(s*stride + n)*2 + {0,1}]
    const float* vin;       // [s*9 + r*3 + c]
    const int* nvis;        // optional per-snapshot counts (device); null: n_const
    size_t stride;
    int n_const;
    int vis_type;
};

// Device tables for one image size.
struct PlanTables {
    int I, G, M;
    float2* tw_g;      // stage twiddles for the length-G transform
    float2* tw_m;      // stage twiddles for the length-M transform
    float2* wpost;     // exp(+2 pi i k / G), k < M (C2R split)
    float* corr;       // 1 / (c(t) * norm), t = 0..I/2
};

// Work buffers for a batch of up to S snapshots of one shape.
struct WorkView {
    VisRec* rec;       // [S*stride]
    int2* bkslot;      // [S*stride]
    VisRec* sorted;    // [S*stride]
    int* hist;         // [S*G]
    int* bstart;       // [S*(G+1)]
    SnapMeta* meta;    // [S]
    SnapGeom* geom;    // [S]
    float2* kt;        // [S][I/kKTile][M+1][kKTile]
    float* pre;        // [S][I][I] corrected image before re-projection
    float* img;        // [S][I][I] re-projected images
    unsigned* work;    // [1] persistent-kernel work counter
    unsigned* pmax;    // [pmax_words(S)] spread slots of the image maxima
    size_t stride;
};

// Tunable launch parameters (the autotuner picks among these).
struct Tuning {
    int stage1_records;   // gridding records per thread held in registers (2, 4, 8)
    int stage2_rows;    // image rows per stage-2 workgroup (1, 2, 4)
    int splat_tile;     // target tile edge of the splat (32 or 64)
    int persistent;     // stage 1 as a persistent work-queue kernel
};

// Stage ids for timing.
enum Stage {
    kStReset = 0,
    kStBin,
    kStScan,
    kStScatter,
    kStGridFft,
    kStRowC2R,
    kStSplat,
    kStSplatTrig,
    kStTrigger,
    kStCount
};
const char* stage_name(int st);

// Kernel launchers (fip_kernels.hip).  All launch asynchronously on stream.
hipError_t launch_reset(const InputView& in, const WorkView& w, const PlanTables& p, int S,
                        float cell, const int* norm_n_host, int norm_n_const, hipStream_t st);
hipError_t launch_bin(const InputView& in, const WorkView& w, const PlanTables& p, int S,
                      int nmax, double scale, hipStream_t st);
hipError_t launch_scan(const WorkView& w, const PlanTables& p, int S, hipStream_t st);
hipError_t launch_scatter(const WorkView& w, const PlanTables& p, int S, int nmax, hipStream_t st);
hipError_t launch_stage1(const WorkView& w, const PlanTables& p, int S, const Tuning& tu,
                         hipStream_t st);
hipError_t launch_stage2(const WorkView& w, const PlanTables& p, int S, const Tuning& tu,
                         hipStream_t st);
// Splat of snapshots 0..S-1 of the view.
hipError_t launch_splat(const WorkView& w, const PlanTables& p, int S, const Tuning& tu,
                        hipStream_t st);

// Fused tail of one window (a, b, s3) with a 32-pixel trigger tile: the
// splat of snapshot s3 of the view computes the trigger terms from the
// window's first two images img_a, img_b (already splatted) and its own
// values, and stores its own image only if store; then the finish pass turns
// the terms into the tLISI.  The same bits as launch_splat of s3 followed by
// launch_trigger.  Fusable when the splat uses 64-pixel tiles and the
// trigger tile is 32.  The terms live in w.kt, which the row pass is done with.
bool splat_trigger_fusable(const PlanTables& p, const Tuning& tu, int tile);
hipError_t launch_splat_trigger(const WorkView& w, const PlanTables& p, int s3, const float* img_a,
                                const float* img_b, bool store, hipStream_t st);
hipError_t launch_trigger_finish(const WorkView& w, const PlanTables& p, int s1, int s2, int s3, int tile, float* out,
                                 hipStream_t st);

// Trigger over nwin windows; win: device int[3*nwin] of snapshot slots in img.
hipError_t launch_trigger(const WorkView& w, const PlanTables& p, const int* win, int nwin,
                          int tile, float* out, hipStream_t st);

// Trigger over the consecutive windows (s0+k, s0+k+1, s0+k+2), k < nwin,
// reading each image once per group of windows.
hipError_t launch_trigger_run(const WorkView& w, const PlanTables& p, int s0, int nwin, int tile, float* out,
                              hipStream_t st);

// meta[s].maxp_bits / maxd_bits from the spread slots, for host readers of
// imaging results without a trigger launch.
hipError_t launch_max_finalize(const WorkView& w, int S, hipStream_t st);

// Inverse FFT of `count` independent length-n arrays in place (self-check).
hipError_t launch_fft_check(int n, float2* data, const float2* tw, int count, hipStream_t st);

// Sizes supported by the compiled FFT plans.
bool fft_size_supported(int n);
// Host generation of the stage twiddle table of the compiled plan for n
// (FP64 -> FP32).  out must hold n-1 entries.
void make_stage_twiddles(int n, float2* out);
// Threads per transform of the compiled plan for n.
int fft_threads(int n);

}  // namespace fip

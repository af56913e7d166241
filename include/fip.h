/*
 * fip.h - C ABI of the MI355X-native FIP-TOI imaging-and-trigger path.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Every
 * function returns a fip_status (0 on success, negative on failure) and
 * leaves a message for fip_last_error() on failure.
 *
 * Conventions shared with the reference (FIPkernels.cu:554-557):
 *   vis_re, vis_im : [snapshot][baseline]            (float unless noted)
 *   bin            : [snapshot][baseline][2]          grid coords in wavelengths
 *   vin            : [snapshot][3][3]                 rows V0, V1, V0 x V1
 *   result         : [tile_row][tile_col], (image/tile)^2 floats
 *   images         : [row][col], image*image floats per snapshot
 */
#ifndef FIP_H
#define FIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fip_status {
    FIP_OK = 0,
    FIP_ERR_ARG = -1,        /* null pointer, zero or non-finite argument      */
    FIP_ERR_NSNAP = -2,      /* fewer than 3 snapshots                         */
    FIP_ERR_SIZE = -3,       /* image size / tile / baseline count unsupported */
    FIP_ERR_ALLOC = -4,      /* device or pinned allocation failed             */
    FIP_ERR_HIP = -5,        /* HIP runtime error                              */
    FIP_ERR_NODEVICE = -6,   /* no usable GPU                                  */
    FIP_ERR_BUDGET = -7,     /* serving: snapshot cannot be admitted           */
    FIP_ERR_IO = -8,         /* FITS or file error                             */
    FIP_ERR_DEGENERATE = -9  /* all images zero: tLISI divides by zero (the
                                result is filled with NaN, as the reference's
                                division by maxall^2 = 0 yields)               */
};

enum fip_vis_type { FIP_VIS_F32 = 0, FIP_VIS_BF16 = 1, FIP_VIS_F16 = 2 };

/* Library identification and diagnostics. */
const char *fip_version(void);
const char *fip_last_error(void);
const char *fip_status_string(int status);
/* Number of visible GPUs (0 when none; never initialises a context). */
int fip_device_count(void);
/* Grid size the path uses for an image size: ceil(1.5*image)
 * (FIPkernels.cu:583); 0 when the size is unsupported. */
size_t fip_grid_size(size_t image_size);

/*
 * Drop-in for FIpipe (FIPkernels.cu:554-862; prototype pipe_interface.cpp:8).
 * Host pointers owned by the caller; inputs are only read.  As in the
 * reference only snapshots 0, 1 and 2 are imaged and the trigger compares
 * them.  Synchronous.  Requires num_snapshots >= 3, image_size a multiple
 * of 32 whose grid 1.5*image is 3*2^k (64..8192), 1 <= unit_size <= image;
 * tiles cover the top-left (image/unit)*unit square as in tlisi
 * (FIPkernels.cu:813, 838).
 */
int fip_pipe(const float *vis_re, const float *vis_im, const float *bin,
             const float *vin, float *result, size_t num_baselines,
             size_t image_size, size_t num_snapshots, float cell_size,
             size_t unit_size);

/*
 * fip_pipe with every array already in device memory (no PCIe crossing).
 * Asynchronous on `stream` (a hipStream_t; NULL selects the library's own
 * stream).  The call returns once the work is enqueued; the result is ready
 * when the stream is.  Calls on one device share work buffers and must not
 * overlap.
 */
int fip_pipe_device(const float *vis_re, const float *vis_im, const float *bin,
                    const float *vin, float *result, size_t num_baselines,
                    size_t image_size, size_t num_snapshots, float cell_size,
                    size_t unit_size, void *stream);

/*
 * fip_pipe_device that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * stream capture, serving loops or harnesses that forbid hipMalloc).  Every
 * intermediate buffer and plan table lives in one static 64 MiB device array
 * per GPU, held until the process ends; the call makes no hipMalloc,
 * hipFree or other allocating call.  Image sizes 64..1024 whose work fits
 * the array (1024^2 with up to 130816 baselines); larger shapes return
 * FIP_ERR_SIZE.  The first call on a device writes the plan tables into the
 * array with synchronous copies; later calls only enqueue kernels on
 * `stream` (NULL: the default stream).  All calls on one GPU share the
 * array: calls that overlap in time on one GPU race.
 */
int fip_pipe_device_noalloc(const float *vis_re, const float *vis_im,
                            const float *bin, const float *vin, float *result,
                            size_t num_baselines, size_t image_size,
                            size_t num_snapshots, float cell_size,
                            size_t unit_size, void *stream);

/*
 * Precision form of fip_pipe: visibilities stored as bfloat16 or IEEE half
 * (vis_type FIP_VIS_BF16 / FIP_VIS_F16; 2-byte elements), converted to FP32
 * on load; gridding accumulates contributions scaled by a power of two per
 * block of rows and rounded to integers in int32 (exact, so independent of
 * summation order) and the FFTs run in FP32 (FIPkernels.cu:188-245 for the
 * arithmetic it replaces).
 */
int fip_pipe_lowp(const void *vis_re, const void *vis_im, int vis_type,
                  const float *bin, const float *vin, float *result,
                  size_t num_baselines, size_t image_size, size_t num_snapshots,
                  float cell_size, size_t unit_size);

/*
 * Imaging half of FIpipe for every snapshot (FIPkernels.cu:645-770):
 * out_pre (corrected image before re-projection, the reference's dirty_pre),
 * out_img (re-projected image dirty_s) and out_max (maxall[s]); any of the
 * three may be NULL.  Host memory, nsnap*image*image floats per image array.
 */
int fip_image(const float *vis_re, const float *vis_im, const float *bin,
              const float *vin, size_t num_baselines, size_t image_size,
              size_t num_snapshots, float cell_size, float *out_pre,
              float *out_img, float *out_max);

/*
 * Trigger half of FIpipe (FIPkernels.cu:806-851): three host images and
 * their maxima -> tLISI.  Returns FIP_ERR_DEGENERATE (result NaN) when all
 * three maxima are zero.
 */
int fip_trigger(const float *d1, const float *d2, const float *d3,
                const float maxall[3], size_t image_size, size_t unit_size,
                float *result);

/*
 * Batching form: nsnap snapshots resident in device memory, each with its
 * own visibility count nvis[s] (host array; rows beyond nvis[s] are
 * ignored), laid out at a common stride (elements between snapshots).  The
 * splat normalises by norm_n (the padded baseline count, as the reference
 * does for zero-padded snapshots, FIPkernels.cu:391; 0 selects nvis[s]).
 * Images every snapshot, then tLISI for every window (s, s+1, s+2):
 * d_tlisi holds (nsnap-2)*(image/tile)^2 floats.  The launch sequence is
 * captured once into a hipGraph per (shape, pointers) and replayed.
 */
int fip_batch_device(const void *vis_re, const void *vis_im, int vis_type,
                     const float *bin, const float *vin, const int *nvis,
                     size_t stride, size_t nsnap, size_t norm_n,
                     size_t image_size, float cell_size, size_t unit_size,
                     float *d_tlisi, void *stream);

/*
 * fip_batch_device with host arrays: copies the nsnap snapshots to the
 * device, images them and returns (nsnap-2)*(image/tile)^2 tLISI values in
 * host memory (window k compares snapshots k, k+1, k+2).  Used by the CLI
 * for combined-FITS input with more than three snapshots.
 */
int fip_batch_host(const void *vis_re, const void *vis_im, int vis_type,
                   const float *bin, const float *vin, const int *nvis,
                   size_t stride, size_t nsnap, size_t norm_n,
                   size_t image_size, float cell_size, size_t unit_size,
                   float *tlisi);

/*
 * SVD-plane uv projection (SVD_MFS.py:38-110; SVDPreprocessKernels.cu:41-206,
 * 320-424).  uvw [nrows][3] in metres, freqs [nchan] in Hz; samples are
 * channel-major (n = row + chan*nrows).  Covariance and eigenvectors in FP64,
 * B = D.[V0 V1] on the GPU as a streaming kernel.  Outputs bin [nsamp][2]
 * and vin [3][3] as float (the FITS interchange type); host memory.
 */
int fip_project(const double *uvw, size_t nrows, const double *freqs,
                size_t nchan, float *bin, float *vin);

/*
 * Raw per-snapshot inputs in device memory for fip_pipe_raw_device (the
 * device-resident form of SVDPreprocessGpu.h:10-32's HostMeasurementSetData,
 * one block per snapshot).  Sample n of a snapshot is row n % nrows of
 * channel n / nrows (channel-major, SVD_MFS.py:48-52, 94-95).
 */
typedef struct fip_raw_view {
    const double *uvw;            /* [nsnap][nrows][3] metres                */
    const double *freqs;          /* [nchan] Hz, shared by every snapshot    */
    const float *xx_re, *xx_im;   /* [nsnap][nchan*nrows] polarisation 0     */
    const float *yy_re, *yy_im;   /* [nsnap][nchan*nrows] polarisation 3     */
    const unsigned char *flag0;   /* [nsnap][nchan*nrows] 0/1                */
    const unsigned char *flag3;
    const float *w0, *w3;         /* [nsnap][nchan*nrows] WEIGHT_SPECTRUM    */
    size_t nrows, nchan, nsnap;
} fip_raw_view;

/*
 * The whole chain on the device, from raw visibilities to trigger: the
 * weighted XX/YY collapse (SVDPreprocessKernels.cu:208-237), the SVD-plane
 * projection B = D.[V0 V1] (SVD_MFS.py:38-75, 98-110;
 * SVDPreprocessKernels.cu:255-424; FP64 moments in one pass, the 3x3
 * eigen-solve on the host after one small read-back), then FIpipe's imaging
 * and trigger (FIPkernels.cu:554-862) of snapshots 0..2 with N = nchan*nrows
 * samples each.  result: device (image/unit)^2 floats.  bin_out ([3][N][2])
 * and vin_out ([3][9]) receive B and V when not NULL (device).  Returns once
 * the work after the eigen-solve is enqueued on `stream` (NULL: the
 * library's stream); the result is ready when the stream is.
 */
int fip_pipe_raw_device(const fip_raw_view *raw, size_t image_size,
                        float cell_size, size_t unit_size, float *result,
                        float *bin_out, float *vin_out, void *stream);

/*
 * Weighted XX/YY collapse (SVD_MFS.py:77-96; SVDPreprocessKernels.cu:208-237):
 * (w0*XX + w3*YY)/(w0+w3) with w = weight*!flag, 0 where w0+w3 <= 0.
 * Host arrays of n samples.
 */
int fip_collapse(const float *xx_re, const float *xx_im, const float *yy_re,
                 const float *yy_im, const unsigned char *flag0,
                 const unsigned char *flag3, const float *w0, const float *w3,
                 size_t n, float *out_re, float *out_im);

/*
 * Measurement Set ingest from exported table columns (SVDPreprocessMain.cpp:
 * 337-546, 548-640, without the casacore reader: the columns come from an
 * external converter such as python-casacore's getcol).
 *   file_index [nrows]: table (MS file) of each row, or NULL (one table);
 *   time [nrows]: TIME (needed with group_by_time);
 *   uvw [nrows][3] metres; data [nrows][nchan_data][npol][2] float (re, im);
 *   flag [nrows][nchan_data][npol] (0/1); weight [nrows][nchan_data][npol]
 *   (WEIGHT_SPECTRUM) or NULL (all 1); freqs [nchan] Hz (CHAN_FREQ of every
 *   SPECTRAL_WINDOW row, concatenated).
 * Snapshots: tables in ascending file_index, each one snapshot or, with
 * group_by_time, one per distinct TIME in ascending order (rows kept in row
 * order).  Per snapshot, XX = pol 0 and YY = pol 3 of channel c read DATA
 * channel c % nchan_data; samples are channel-major; UVW and frequencies are
 * rounded to float as the reference's reader stores them; then fip_collapse
 * and fip_project.  `workers` snapshots are preprocessed concurrently (0:
 * min(4, hardware threads, snapshots)).  Snapshots are zero-padded to the
 * largest sample count.  Writes <prefix>Bin.fits, Vin.fits, Visreal.fits,
 * Visimag.fits and the 3-extension combined file `combined` (NULL or "":
 * <prefix>input.fits).  *nsnap_out / *nsamp_out (optional) get the snapshot
 * count and padded samples per snapshot.  FIP_ERR_ARG for npol < 4.
 */
int fip_ingest_table(const int *file_index, const double *time,
                     const double *uvw, size_t nrows, const float *data,
                     const unsigned char *flag, const float *weight,
                     size_t nchan_data, size_t npol, const double *freqs,
                     size_t nchan, int group_by_time, int workers,
                     const char *prefix, const char *combined,
                     size_t *nsnap_out, size_t *nsamp_out);

/* ---- per-stage timing (hipEvents on the launching stream) -------------- */
/* Enable or disable per-stage event timing for subsequent calls. */
int fip_set_profiling(int on);
/* Stage names and accumulated milliseconds / launch counts since the last
 * reset; n_stages <= fip_stage_count(). */
int fip_stage_count(void);
const char *fip_stage_name(int stage);
int fip_stage_times(double *ms, long *counts, int n_stages);
int fip_stage_times_reset(void);

/* ---- launch tuning and autotuner ---------------------------------------- */
/* Measure candidate launch configurations (gridding records per thread,
 * rows per C2R workgroup, splat tile, persistent gridding) on synthetic data
 * of the given shape, keep the fastest for that image size, and write it to
 * *choice (4 ints: records, rows, tile, persistent; may be NULL). */
int fip_autotune(size_t image_size, size_t num_baselines, float cell_size,
                 int reps, int *choice);
/* Force a launch configuration for an image size (choice as above). */
int fip_set_tuning(size_t image_size, const int *choice);

/* ---- serving form -------------------------------------------------------- */
typedef struct fip_server fip_server;
/* One entry per sub-array stream. */
typedef struct {
    size_t num_baselines;
    size_t image_size;
    float cell_size;
    size_t unit_size;
    int vis_type;           /* fip_vis_type of submitted visibilities */
} fip_stream_config;
/* Create a server with nstreams sub-array streams, a device ring of `ring`
 * image slots per stream and an HBM budget in bytes for all device state.
 * The streams' fixed state (work buffers, image ring) plus one snapshot of
 * the largest stream's inputs must fit (else FIP_ERR_BUDGET); the rest of
 * the budget, up to `ring` snapshots per stream, is an input pool shared by
 * all streams, from which each arrival reserves its inputs. */
int fip_server_create(const fip_stream_config *streams, int nstreams, int ring,
                      size_t hbm_budget_bytes, fip_server **out);
/* Submit one snapshot (host arrays) for stream `sid` with its arrival time
 * in seconds of the monotonic clock (CLOCK_MONOTONIC, std::steady_clock,
 * Python's time.monotonic()); a negative value stamps the submit time.
 * Admission per arrival: returns FIP_ERR_BUDGET (and counts the snapshot
 * as refused) when its inputs do not fit the free part of the input pool;
 * the reservation is released when the snapshot's tLISI is on the host. */
int fip_server_submit(fip_server *srv, int sid, const void *vis_re,
                      const void *vis_im, const float *bin, const float *vin,
                      double arrival_s);
/* Block until every submitted snapshot is processed. */
int fip_server_drain(fip_server *srv);
/* Latency statistics (ms, arrival -> tLISI on host) and counts. */
int fip_server_stats(fip_server *srv, double *p50_ms, double *p99_ms,
                     double *mean_ms, long *processed, long *refused);
/* Launch sequences enqueued so far, summed over the streams.  A snapshot
 * submitted while its stream's previous sequence is still running waits and
 * is launched with the others queued behind it (up to 3 per sequence, fewer
 * where the ring wraps), so processed / sequences > 1 under load. */
int fip_server_sequences(fip_server *srv, long *sequences);
/* Latest tLISI matrix of a stream ((image/tile)^2 floats) and its window
 * index; FIP_ERR_ARG if the stream has fewer than 3 images yet. */
int fip_server_latest(fip_server *srv, int sid, float *tlisi, long *window);
void fip_server_destroy(fip_server *srv);

/* ---- synthetic inputs (BASELINE workload generator) ---------------------- */
/* Log-radial antenna layout, earth-rotation uvw, FP64 DFT visibilities of a
 * power-law point-source sky plus transients, Gaussian noise, SVD-plane
 * projection.  Host outputs for nsnap snapshots of nbl = nant*(nant-1)/2
 * baselines: vis_re/vis_im [nsnap][nbl], bin [nsnap][nbl][2],
 * vin [nsnap][9].  transient_snap[k] (k < ntransient) is the snapshot in
 * which transient k is on; flag_frac_max > 0 zeroes a random fraction in
 * [0, flag_frac_max) of each snapshot's visibilities and reports the
 * surviving count in nvis_out (may be NULL). */
int fip_synth(int nant, double radius_m, size_t nsnap, double dt_s,
              double freq_hz, double dec_deg, double lat_deg, size_t image_size,
              float cell, int nsrc, double smin, double smax, double index,
              double noise_sigma, int ntransient, const int *transient_snap,
              double transient_flux, double flag_frac_max,
              unsigned long long seed, float *vis_re, float *vis_im, float *bin,
              float *vin, int *nvis_out);

/* ---- FITS I/O (pipe_interface.cpp:12-137) --------------------------------- */
/* Read the primary image of a FITS file as float (any BITPIX, NAXIS 1..3).
 * *data is malloc'd; dims[0..2] get NAXIS1..3 (1 for absent axes). */
int fip_fits_read(const char *path, float **data, long dims[3]);
/* Read an image extension by EXTNAME (or the primary HDU when extname is
 * NULL or ""). */
int fip_fits_read_ext(const char *path, const char *extname, float **data,
                      long dims[3]);
/* Write a 2-D BITPIX -32 primary image (replacing an existing file). */
int fip_fits_write(const char *path, const float *data, long naxis1,
                   long naxis2);
/* Write a 3-D BITPIX -32 primary image (e.g. one tLISI matrix per window). */
int fip_fits_write3(const char *path, const float *data, long naxis1,
                    long naxis2, long naxis3);
void fip_free(void *p);

/* ---- internal checks ----------------------------------------------------- */
/* Run the LDS FFT kernel (inverse, unnormalised) on `count` independent
 * length-n complex arrays (host, interleaved re/im, in place). */
int fip_debug_fft(int n, float *data, int count);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
/* C++-linkage drop-in with the reference's exact signature
 * (pipe_interface.cpp:8; mangled _Z6FIpipePfS_S_S_S_mmmfm); forwards to
 * fip_pipe and returns its status. */
int FIpipe(float *Visreal, float *Visimag, float *Bin, float *Vin,
           float *result_array, size_t num_baselines, size_t image_size,
           size_t num_snapshots, float cell_size, size_t unit_size);
#endif

#endif /* FIP_H */

/*
 * FP64 CPU oracle for the FIP-TOI imaging-and-trigger path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as the product.  The product path lives
 * in fastimagingpipe_amd/ and fails loudly when its HIP library is missing.
 *
 * Every function restates one piece of the reference algorithm in double
 * precision and cites the reference file:line it follows.  How each function
 * is pinned (closed forms, numpy/sklearn cross-checks, brute force) is listed
 * in the header comment of oracle/fip_oracle.c.
 *
 * Array conventions (identical to the reference's FIpipe arguments):
 *   vis_re/vis_im : [snapshot][baseline]          float
 *   bin           : [snapshot][baseline][2]       float, wavelengths
 *   vin           : [snapshot][3][3]              float, rows V0, V1, V0xV1
 *   images        : [row][col], row-major, double
 *   result        : [tile_row][tile_col]
 */
#ifndef FIP_ORACLE_H
#define FIP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ES kernel exp(beta*(sqrt(1-x^2)-1)), 0 outside |x|<=1 (FIPkernels.cu:67-71). */
double fo_es(double x);
/* Positive half of 28-point Gauss-Legendre, nodes descending, and ES(node)
 * (the tables at FIPkernels.cu:14-29). */
void fo_quadrature_tables(double nodes[14], double weights[14], double kern[14]);
/* Grid size ceil(1.5*I) (FIPkernels.cu:583). */
size_t fo_grid_size(size_t image);
/* Grid-correction function c(t), t = 0..image/2 (FIPkernels.cu:173-186). */
void fo_conv_corr(size_t image, size_t grid, double *c);

/* In-place complex FFT of length n (interleaved re,im), any n >= 1.
 * sign=+1 gives the unnormalised inverse sum x[j]*exp(+2*pi*i*j*k/n). */
int fo_fft(double *data, size_t n, int sign);
/* In-place 2-D inverse FFT of a g x g complex grid (cuFFT C2C INVERSE
 * semantics, FIPkernels.cu:635-638, 679).  nthreads <= 0 means default. */
int fo_fft2_inverse(double *grid, size_t g, int nthreads);

/* computeVisWeighted + gridding for one snapshot (FIPkernels.cu:188-245).
 * grid: g*g complex (interleaved), accumulated (caller zeroes it). */
void fo_grid(const float *vis_re, const float *vis_im, const float *bin,
             const float *vin, size_t nbl, size_t g, float cell, double *grid);
/* accumulation + scaling: crop, (-1)^(|x|+|y|), /(c c norm^2), fabs
 * (FIPkernels.cu:256-287).  dirty_pre: I*I. */
void fo_crop_correct(const double *fgrid, const double *c, size_t image,
                     size_t g, double *dirty_pre);
/* coordschange (FIPkernels.cu:289-299): out_index[(i*I+j)*2+{0,1}]. */
void fo_coordschange(const float *vin, size_t image, double *out_index);
/* p2p SIN re-projection by the degree-trig chain (FIPkernels.cu:301-376,
 * helpers :73-171), in place on out_index. */
void fo_p2p(double *out_index, const float *vin, float cell, size_t image);
/* finalinterp bilinear forward splat (FIPkernels.cu:378-413), accumulates
 * into dirty (caller zeroes it). */
void fo_splat(const double *out_index, const double *dirty_pre, size_t image,
              size_t nbl, double *dirty);
/* max_large + max_small: max(0, max pixel) (FIPkernels.cu:431-497). */
double fo_max(const double *img, size_t npix);
/* FI trigger: subtraction x3, setNonPositiveToC, tlisi (FIPkernels.cu:806-851,
 * 415-429, 499-552).  result: (I/T)^2. */
void fo_trigger(const double *d1, const double *d2, const double *d3,
                const double maxall[3], size_t image, size_t tile,
                double *result);

/* Whole imaging stage for one snapshot: weighting, gridding, FFT, crop,
 * correction, map, splat and max (FIPkernels.cu:647-761 per snapshot).
 * dirty_pre (may be NULL) and dirty: I*I.  Returns 0 or negative on error. */
int fo_image(const float *vis_re, const float *vis_im, const float *bin,
             const float *vin, size_t nbl, size_t image, float cell,
             double *dirty_pre, double *dirty, double *maxval, int nthreads);
/* FIpipe (FIPkernels.cu:554-862): the first three snapshots -> tLISI. */
int fo_pipe(const float *vis_re, const float *vis_im, const float *bin,
            const float *vin, double *result, size_t nbl, size_t image,
            size_t nsnap, float cell, size_t tile, int nthreads);

/* SVD_MFS.py:38-110 projection: uvw [nrows][3] metres, freqs [nchan] Hz ->
 * samples n = row + chan*nrows; bin [nsamp][2], vin [3][3] (float64). */
int fo_project(const double *uvw, size_t nrows, const double *freqs,
               size_t nchan, double *bin, double *vin);
/* SVD_MFS.py:77-96 weighted XX/YY collapse for one sample array. */
void fo_collapse(const double *xx_re, const double *xx_im, const double *yy_re,
                 const double *yy_im, const unsigned char *flag0,
                 const unsigned char *flag3, const double *w0, const double *w3,
                 size_t n, double *out_re, double *out_im);

#ifdef __cplusplus
}
#endif
#endif

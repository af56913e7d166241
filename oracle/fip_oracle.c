/*
 * FP64 CPU oracle for the FIP-TOI imaging-and-trigger path (see fip_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  Nothing in fastimagingpipe_amd/ links or calls it.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   fo_es, fo_quadrature_tables : against the reference's tabulated nodes,
 *       weights and kernel values (FIPkernels.cu:14-29, fixture
 *       tests/golden/reference_quadrature.json) and numpy's leggauss(28).
 *   fo_conv_corr   : c(0) = 1 (closed form: 4*int ES = norm), and against a
 *       direct numerical cosine transform of ES (scipy.integrate.quad).
 *   fo_fft, fo_fft2_inverse : against numpy.fft and a brute-force DFT.
 *   fo_grid + fo_fft2_inverse + fo_crop_correct : the one-visibility known
 *       answer |Re(v e^{2 pi i (p1 y + p2 x)/g})| / J / c / c * (ES(.)...) and
 *       a brute-force direct tap-DFT of the whole chain; linearity.
 *   fo_coordschange + fo_p2p : V = identity gives the column mirror
 *       (2h-i, j); against the closed SIN form P1 - xi n'/dc, P2 + eta n'/dc;
 *       a point source of a flat array seen in random tilted frames lands at
 *       row sign(V22) l/cell + h, column -sign(V22) m/cell + h (sky geometry
 *       and V orthonormal; mutation tests plant a swapped V01/V10, a dropped
 *       or signed |V22| divisor and check that this pin fails).
 *   fo_splat : integer targets copy exactly; mass conservation; a source at
 *       fractional (LL, MM) keeps its centroid and gives a rank-one 2 x 2
 *       footprint (a swapped {LL}/{MM} fails it).
 *   fo_project (end to end): a source placed on pre-image pixel (y, x) of
 *       the PCA frame images to exactly N/|V22| there (un-centred B).
 *   fo_max, fo_trigger : constant images give tLISI = 1; a single changed
 *       pixel gives the hand-computed tile value; numpy brute force.
 *   fo_project : against scikit-learn PCA fixtures (tests/golden/).
 *   fo_collapse : hand-computed cases: equal weights give the mean, one
 *       side flagged gives the other, all flagged or zero weights give 0.
 */
#include "fip_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FO_PI 3.14159265358979323846
#define FO_BETA 15.3704324328      /* FIPkernels.cu:205 */
#define FO_SUPPORT 8               /* FIPkernels.cu:174, 198 */
#define FO_WCSTRIG_TOL 1e-10       /* FIPkernels.cu:10 */

/* ------------------------------------------------------------------ */
/* ES kernel and quadrature tables                                     */
/* ------------------------------------------------------------------ */

double fo_es(double x)
{
    /* exp_semicircle, FIPkernels.cu:67-71 */
    const double xx = x * x;
    return (xx > 1.0) ? 0.0 : exp(FO_BETA * (sqrt(1.0 - xx) - 1.0));
}

/* Legendre P_n(x) and its derivative by the three-term recurrence. */
static void legendre(int n, double x, double *p, double *dp)
{
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    *p = p1;
    *dp = n * (x * p1 - p0) / (x * x - 1.0);
}

void fo_quadrature_tables(double nodes[14], double weights[14], double kern[14])
{
    /* FIPkernels.cu:14-29: positive half of 28-point Gauss-Legendre in
     * descending node order, and the ES kernel at those nodes. */
    const int n = 28;
    for (int i = 0; i < 14; ++i) {
        double x = cos(FO_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre(n, x, &p, &dp);
            double dx = p / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        double p, dp;
        legendre(n, x, &p, &dp);
        nodes[i] = x;
        weights[i] = 2.0 / ((1.0 - x * x) * dp * dp);
        kern[i] = fo_es(x);
    }
}

/* 4 * integral_{-1}^{1} ES(x) dx, by 400-panel Gauss-Legendre (the reference
 * hard-codes this constant, FIPkernels.cu:585). */
static double es_norm(void)
{
    static double cached = 0.0;
    if (cached != 0.0) return cached;
    double nd[14], wt[14], kn[14];
    fo_quadrature_tables(nd, wt, kn);
    const int panels = 400;
    double acc = 0.0;
    for (int p = 0; p < panels; ++p) {
        double a = -1.0 + 2.0 * p / panels, b = -1.0 + 2.0 * (p + 1) / panels;
        double mid = 0.5 * (a + b), half = 0.5 * (b - a);
        for (int i = 0; i < 14; ++i) {
            acc += half * wt[i] * (fo_es(mid + half * nd[i]) + fo_es(mid - half * nd[i]));
        }
    }
    cached = 4.0 * acc;
    return cached;
}

size_t fo_grid_size(size_t image)
{
    /* computeCeil(1.5*float(I)), FIPkernels.cu:31-37, 583 */
    float num = (float)(1.5 * (double)(float)image);
    return (size_t)ceilf(num);
}

void fo_conv_corr(size_t image, size_t grid, double *c)
{
    /* convolveKernel, FIPkernels.cu:173-186 */
    double nd[14], wt[14], kn[14];
    fo_quadrature_tables(nd, wt, kn);
    const double norm = es_norm();
    for (size_t t = 0; t <= image / 2; ++t) {
        double tn = (double)t / (double)grid;
        double acc = 0.0;
        for (int i = 0; i < 14; ++i) {
            acc += kn[i] * wt[i] * cos(FO_PI * tn * FO_SUPPORT * nd[i]);
        }
        c[t] = acc * FO_SUPPORT / norm;
    }
}

/* ------------------------------------------------------------------ */
/* FP64 mixed-radix Stockham FFT                                       */
/* ------------------------------------------------------------------ */

static int factorize(size_t n, size_t *f)
{
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    for (size_t p = 3; p * p <= n; p += 2) {
        while (n % p == 0) { f[nf++] = p; n /= p; }
    }
    if (n > 1) f[nf++] = n;
    return nf;
}

/* One FFT of length n using scratch (2n doubles) and a root table
 * w[k] = exp(sign*2*pi*i*k/n) (2n doubles). */
static void fft_core(double *a, double *scratch, size_t n, const double *w,
                     const size_t *f, int nf)
{
    double *src = a, *dst = scratch;
    size_t ns = 1;
    double vr[64], vi[64], orr[64], oi[64];
    for (int s = 0; s < nf; ++s) {
        const size_t r = f[s];
        const size_t m = n / r;          /* butterflies per stage */
        const size_t step = n / (ns * r);
        for (size_t j = 0; j < m; ++j) {
            const size_t k = j % ns;
            double *pr = vr, *pi = vi, *qr = orr, *qi = oi;
            double *heap = NULL;
            if (r > 64) {
                heap = (double *)malloc(4 * r * sizeof(double));
                pr = heap; pi = heap + r; qr = heap + 2 * r; qi = heap + 3 * r;
            }
            for (size_t q = 0; q < r; ++q) {
                double xr = src[2 * (j + q * m)], xi = src[2 * (j + q * m) + 1];
                size_t ti = (q * k * step) % n;
                double cr = w[2 * ti], ci = w[2 * ti + 1];
                pr[q] = xr * cr - xi * ci;
                pi[q] = xr * ci + xi * cr;
            }
            for (size_t u = 0; u < r; ++u) {
                double sr = 0.0, si = 0.0;
                for (size_t q = 0; q < r; ++q) {
                    size_t ti = ((q * u) % r) * (n / r);
                    double cr = w[2 * ti], ci = w[2 * ti + 1];
                    sr += pr[q] * cr - pi[q] * ci;
                    si += pr[q] * ci + pi[q] * cr;
                }
                qr[u] = sr;
                qi[u] = si;
            }
            const size_t base = (j / ns) * ns * r + k;
            for (size_t u = 0; u < r; ++u) {
                dst[2 * (base + u * ns)] = qr[u];
                dst[2 * (base + u * ns) + 1] = qi[u];
            }
            free(heap);
        }
        double *t = src; src = dst; dst = t;
        ns *= r;
    }
    if (src != a) memcpy(a, src, 2 * n * sizeof(double));
}

static double *root_table(size_t n, int sign)
{
    double *w = (double *)malloc(2 * n * sizeof(double));
    if (!w) return NULL;
    for (size_t k = 0; k < n; ++k) {
        double ang = 2.0 * FO_PI * (double)k / (double)n;
        w[2 * k] = cos(ang);
        w[2 * k + 1] = sign * sin(ang);
    }
    return w;
}

int fo_fft(double *data, size_t n, int sign)
{
    if (n == 0) return -1;
    size_t f[64];
    int nf = factorize(n, f);
    double *w = root_table(n, sign >= 0 ? 1 : -1);
    double *scratch = (double *)malloc(2 * n * sizeof(double));
    if (!w || !scratch) { free(w); free(scratch); return -1; }
    fft_core(data, scratch, n, w, f, nf);
    free(w);
    free(scratch);
    return 0;
}

int fo_fft2_inverse(double *grid, size_t g, int nthreads)
{
    size_t f[64];
    int nf = factorize(g, f);
    double *w = root_table(g, 1);
    if (!w) return -1;
    int ok = 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    (void)nthreads;
#endif
    /* rows */
#pragma omp parallel num_threads(nthreads)
    {
        double *scratch = (double *)malloc(2 * g * sizeof(double));
        if (!scratch) {
#pragma omp atomic write
            ok = 0;
        } else {
#pragma omp for schedule(static)
            for (long r = 0; r < (long)g; ++r) {
                fft_core(grid + 2 * (size_t)r * g, scratch, g, w, f, nf);
            }
        }
        free(scratch);
    }
    /* columns, 8 at a time through a gathered buffer */
#pragma omp parallel num_threads(nthreads)
    {
        double *col = (double *)malloc(2 * g * sizeof(double));
        double *scratch = (double *)malloc(2 * g * sizeof(double));
        if (!col || !scratch) {
#pragma omp atomic write
            ok = 0;
        } else {
#pragma omp for schedule(static)
            for (long c = 0; c < (long)g; ++c) {
                for (size_t r = 0; r < g; ++r) {
                    col[2 * r] = grid[2 * (r * g + (size_t)c)];
                    col[2 * r + 1] = grid[2 * (r * g + (size_t)c) + 1];
                }
                fft_core(col, scratch, g, w, f, nf);
                for (size_t r = 0; r < g; ++r) {
                    grid[2 * (r * g + (size_t)c)] = col[2 * r];
                    grid[2 * (r * g + (size_t)c) + 1] = col[2 * r + 1];
                }
            }
        }
        free(col);
        free(scratch);
    }
    free(w);
    return ok ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* Imaging                                                             */
/* ------------------------------------------------------------------ */

void fo_grid(const float *vis_re, const float *vis_im, const float *bin,
             const float *vin, size_t nbl, size_t g, float cell, double *grid)
{
    /* computeVisWeighted, FIPkernels.cu:188-195 */
    const double J = fabs((double)vin[0] * vin[4] - (double)vin[1] * vin[3]);
    /* gridding, FIPkernels.cu:197-245; r1r2_scale = cell*grid (:584) */
    const double scale = (double)cell * (double)g;
    const long gmin = -(long)g / 2;
    const long gmax = ((long)g - 1) / 2;
    const long origin = (long)g / 2;
    for (size_t n = 0; n < nbl; ++n) {
        const double vr = (double)vis_re[n] / J;
        const double vi = (double)vis_im[n] / J;
        const double p1 = (double)bin[2 * n] * scale;
        const double p2 = (double)bin[2 * n + 1] * scale;
        long r1lo = (long)ceil(p1 - 4.0), r1hi = (long)floor(p1 + 4.0);
        long r2lo = (long)ceil(p2 - 4.0), r2hi = (long)floor(p2 + 4.0);
        if (r1lo < gmin) r1lo = gmin;
        if (r1hi > gmax) r1hi = gmax;
        if (r2lo < gmin) r2lo = gmin;
        if (r2hi > gmax) r2hi = gmax;
        if (r1lo > r1hi || r2lo > r2hi) continue;
        double k2[16];
        for (long r2 = r2lo; r2 <= r2hi; ++r2) k2[r2 - r2lo] = fo_es(((double)r2 - p2) / 4.0);
        for (long r1 = r1lo; r1 <= r1hi; ++r1) {
            const double k1 = fo_es(((double)r1 - p1) / 4.0);
            for (long r2 = r2lo; r2 <= r2hi; ++r2) {
                double kv = k1 * k2[r2 - r2lo];
                if (((r1 + r2) & 1) != 0) kv = -kv;
                const size_t off = (size_t)((r1 + origin) * (long)g + (r2 + origin));
                grid[2 * off] += vr * kv;
                grid[2 * off + 1] += vi * kv;
            }
        }
    }
}

void fo_crop_correct(const double *fgrid, const double *c, size_t image,
                     size_t g, double *dirty_pre)
{
    /* accumulation (FIPkernels.cu:256-272) + scaling (:274-287) */
    const double norm = es_norm();
    const long h = (long)(image / 2);
    const size_t gc = g * g / 2 + g / 2;
    for (long idy = 0; idy < (long)image; ++idy) {
        for (long idx = 0; idx < (long)image; ++idx) {
            const long x = idx - h, y = idy - h;
            double v = fgrid[2 * (size_t)((long)gc + y * (long)g + x)];
            if (((labs(x) + labs(y)) & 1) != 0) v = -v;
            v *= 1.0 / (c[labs(x)] * c[labs(y)] * norm * norm);
            dirty_pre[(size_t)idy * image + (size_t)idx] = fabs(v);
        }
    }
}

void fo_coordschange(const float *vin, size_t image, double *out_index)
{
    /* coordschange, FIPkernels.cu:289-299 */
    const double h = (double)(image / 2);
    const double a22 = fabs((double)vin[8]);
    for (size_t i = 0; i < image; ++i) {
        for (size_t j = 0; j < image; ++j) {
            const double a = (double)i - h, b = (double)j - h;
            out_index[(i * image + j) * 2 + 0] = (-(double)vin[0] * a + (double)vin[3] * b) / a22 + h;
            out_index[(i * image + j) * 2 + 1] = (-(double)vin[1] * a + (double)vin[4] * b) / a22 + h;
        }
    }
}

/* Degree-argument trig helpers, FIPkernels.cu:73-171. */
static void sincosd(double angle, double *s, double *c)
{
    if (fmod(angle, 90.0) == 0.0) {
        int i = (int)fabs(floor(angle / 90.0 + 0.5)) % 4;
        switch (i) {
        case 0: *s = 0.0; *c = 1.0; break;
        case 1: *s = (angle > 0.0) ? 1.0 : -1.0; *c = 0.0; break;
        case 2: *s = 0.0; *c = -1.0; break;
        default: *s = (angle > 0.0) ? -1.0 : 1.0; *c = 0.0; break;
        }
    } else {
        *s = sin(angle * FO_PI / 180.0);
        *c = cos(angle * FO_PI / 180.0);
    }
}

static double sind(double angle)
{
    if (fmod(angle, 90.0) == 0.0) {
        int i = (int)fabs(floor(angle / 90.0 - 0.5)) % 4;
        static const double t[4] = {1.0, 0.0, -1.0, 0.0};
        return t[i];
    }
    return sin(angle * FO_PI / 180.0);
}

static double cosd(double angle)
{
    if (fmod(angle, 90.0) == 0.0) {
        int i = (int)fabs(floor(angle / 90.0 + 0.5)) % 4;
        static const double t[4] = {1.0, 0.0, -1.0, 0.0};
        return t[i];
    }
    return cos(angle * FO_PI / 180.0);
}

static double atan2d(double y, double x)
{
    if (y == 0.0) return (x >= 0.0) ? 0.0 : 180.0;
    if (x == 0.0) return (y > 0.0) ? 90.0 : -90.0;
    return atan2(y, x) * 180.0 / FO_PI;
}

static double acosd(double v)
{
    if (v >= 1.0 && v - 1.0 < FO_WCSTRIG_TOL) return 0.0;
    if (v == 0.0) return 90.0;
    if (v <= -1.0 && v + 1.0 > -FO_WCSTRIG_TOL) return 180.0;
    return acos(v) * 180.0 / FO_PI;
}

static double asind(double v)
{
    if (v <= -1.0 && v + 1.0 > -FO_WCSTRIG_TOL) return -90.0;
    if (v == 0.0) return 0.0;
    if (v >= 1.0 && v - 1.0 < FO_WCSTRIG_TOL) return 90.0;
    return asin(v) * 180.0 / FO_PI;
}

void fo_p2p(double *out_index, const float *vin, float cell, size_t image)
{
    /* p2p, FIPkernels.cu:301-376 (Calabretta & Greisen 2002 SIN). */
    const double xi = (double)vin[6] / (double)vin[8];
    const double eta = (double)vin[7] / (double)vin[8];
    const double dc = (double)cell / FO_PI * 180.0;
    const double c0 = (double)image / 2.0 + 1.0;
    const size_t npix = image * image;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)npix; ++q) {
        double *o = out_index + 2 * (size_t)q;
        const double p1 = o[0], p2 = o[1];
        double x = -dc * (p1 - c0);
        double y = dc * (p2 - c0);
        double r0 = 180.0 / FO_PI;
        const double x0 = x / r0, y0 = y / r0;
        const double r2 = x0 * x0 + y0 * y0;
        const double phi = (r2 != 0.0) ? atan2d(x0, -y0) : 0.0;
        double theta;
        if (r2 < 0.5) theta = acosd(sqrt(r2));
        else if (r2 <= 1.0) theta = asind(sqrt(1.0 - r2));
        else continue;
        double sphi, cphi;
        sincosd(phi, &sphi, &cphi);
        x = sphi;
        y = cphi;
        const double t = (90.0 - fabs(theta)) / 180.0 * FO_PI;
        double z, costhe;
        if (t < 1.0e-5) {
            z = (theta > 0.0) ? t * t / 2.0 : 2.0 - t * t / 2.0;
            costhe = t;
        } else {
            z = 1.0 - sind(theta);
            costhe = cosd(theta);
        }
        r0 = 180.0 / FO_PI;
        const double r = r0 * costhe;
        const double w = xi * xi + eta * eta;
        if (w == 0.0) {
            x = r * x;
            y = -r * y;
        } else {
            z = z * r0;
            x = r * x + xi * z;
            y = -r * y + eta * z;
        }
        o[0] = -1.0 / dc * x + c0;
        o[1] = 1.0 / dc * y + c0;
    }
}

void fo_splat(const double *out_index, const double *dirty_pre, size_t image,
              size_t nbl, double *dirty)
{
    /* finalinterp, FIPkernels.cu:378-413 */
    const long h = (long)(image / 2);
    const double lim = (double)(image / 2 - 1);
    const double inv = 1.0 / (double)nbl;
    for (long idy = 0; idy < (long)image; ++idy) {
        for (long idx = 0; idx < (long)image; ++idx) {
            const size_t q = (size_t)idx * image + (size_t)idy;
            const double LL = out_index[2 * q] - (double)h;
            const double MM = out_index[2 * q + 1] - (double)h;
            const double val = dirty_pre[(size_t)idy * image + (size_t)idx] * inv;
            if (fabs(LL) < lim && fabs(MM) < lim) {
                const long fl = (long)floor(LL), cl = (long)ceil(LL);
                const long fm = (long)floor(MM), cm = (long)ceil(MM);
                const double wl0 = 1.0 - LL + floor(LL), wl1 = LL - floor(LL);
                const double wm0 = 1.0 - MM + floor(MM), wm1 = MM - floor(MM);
                dirty[(size_t)((fm + h) * (long)image + fl + h)] += wl0 * wm0 * val;
                dirty[(size_t)((cm + h) * (long)image + fl + h)] += wl0 * wm1 * val;
                dirty[(size_t)((fm + h) * (long)image + cl + h)] += wl1 * wm0 * val;
                dirty[(size_t)((cm + h) * (long)image + cl + h)] += wl1 * wm1 * val;
            }
        }
    }
}

double fo_max(const double *img, size_t npix)
{
    /* max_large / max_small start their reductions from 0 (FIPkernels.cu:450, 461) */
    double m = 0.0;
    for (size_t i = 0; i < npix; ++i) if (img[i] > m) m = img[i];
    return m;
}

void fo_trigger(const double *d1, const double *d2, const double *d3,
                const double maxall[3], size_t image, size_t tile,
                double *result)
{
    /* FIPkernels.cu:806-851 with subtraction :423-429, setNonPositiveToC
     * :415-421 (C = 1e-6f, :811) and tlisi :499-552. */
    const double C = (double)1e-6f;
    const size_t un = image / tile;
    const double mx3 = fmax(maxall[0], fmax(maxall[1], maxall[2]));
    const double T = (double)tile;
    for (size_t ti = 0; ti < un; ++ti) {
        for (size_t tj = 0; tj < un; ++tj) {
            double sum = 0.0, mx = 0.0, sr = 0.0;
            for (size_t r = 0; r < tile; ++r) {
                for (size_t cidx = 0; cidx < tile; ++cidx) {
                    const size_t p = (ti * tile + r) * image + tj * tile + cidx;
                    const double dd = fabs(fabs(d1[p] - d2[p]) - fabs(d2[p] - d3[p]));
                    const double s2 = (d2[p] <= 0.0) ? C : d2[p];
                    sum += dd;
                    if (dd > mx) mx = dd;
                    const double ratio = dd / s2;
                    sr += (ratio < 1.0) ? ratio : 1.0;
                }
            }
            result[ti * un + tj] = 1.0 - (sum / T / T) * mx * (sr / T / T) / mx3 / mx3;
        }
    }
}

int fo_image(const float *vis_re, const float *vis_im, const float *bin,
             const float *vin, size_t nbl, size_t image, float cell,
             double *dirty_pre, double *dirty, double *maxval, int nthreads)
{
    const size_t g = fo_grid_size(image);
    if (image < 4 || (g & 1) || nbl == 0) return -1;
    double *grid = (double *)calloc(2 * g * g, sizeof(double));
    double *c = (double *)malloc((image / 2 + 1) * sizeof(double));
    double *pre = dirty_pre ? dirty_pre : (double *)malloc(image * image * sizeof(double));
    double *idx = (double *)malloc(2 * image * image * sizeof(double));
    if (!grid || !c || !pre || !idx) {
        free(grid); free(c); free(idx);
        if (!dirty_pre) free(pre);
        return -2;
    }
    fo_conv_corr(image, g, c);
    fo_grid(vis_re, vis_im, bin, vin, nbl, g, cell, grid);
    int rc = fo_fft2_inverse(grid, g, nthreads);
    if (rc == 0) {
        fo_crop_correct(grid, c, image, g, pre);
        fo_coordschange(vin, image, idx);
        fo_p2p(idx, vin, cell, image);
        memset(dirty, 0, image * image * sizeof(double));
        fo_splat(idx, pre, image, nbl, dirty);
        if (maxval) *maxval = fo_max(dirty, image * image);
    }
    free(grid); free(c); free(idx);
    if (!dirty_pre) free(pre);
    return rc;
}

int fo_pipe(const float *vis_re, const float *vis_im, const float *bin,
            const float *vin, double *result, size_t nbl, size_t image,
            size_t nsnap, float cell, size_t tile, int nthreads)
{
    if (nsnap < 3 || tile == 0 || tile > image) return -1;
    const size_t np = image * image;
    double *d = (double *)malloc(3 * np * sizeof(double));
    if (!d) return -2;
    double mx[3];
    for (int s = 0; s < 3; ++s) {
        int rc = fo_image(vis_re + (size_t)s * nbl, vis_im + (size_t)s * nbl,
                          bin + (size_t)s * nbl * 2, vin + (size_t)s * 9, nbl,
                          image, cell, NULL, d + (size_t)s * np, &mx[s], nthreads);
        if (rc) { free(d); return rc; }
    }
    fo_trigger(d, d + np, d + 2 * np, mx, image, tile, result);
    free(d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Preprocessing (SVD_MFS.py)                                          */
/* ------------------------------------------------------------------ */

/* Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix. */
static void jacobi3(double a[3][3], double evec[3][3], double eval[3])
{
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off == 0.0) break;
        for (int p = 0; p < 2; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; ++k) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = cs * akp - sn * akq;
                    a[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = cs * apk - sn * aqk;
                    a[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = cs * vkp - sn * vkq;
                    v[k][q] = sn * vkp + cs * vkq;
                }
            }
        }
    }
    for (int i = 0; i < 3; ++i) {
        eval[i] = a[i][i];
        for (int k = 0; k < 3; ++k) evec[k][i] = v[k][i];   /* column i */
    }
}

int fo_project(const double *uvw, size_t nrows, const double *freqs,
               size_t nchan, double *bin, double *vin)
{
    /* SVD_MFS.py:38-52: baselines scaled by nu/C, flattened channel-major. */
    const double C = 299792458.0;
    const size_t ns = nrows * nchan;
    if (ns < 2) return -1;
    double mu[3] = {0, 0, 0};
    for (size_t ch = 0; ch < nchan; ++ch)
        for (size_t r = 0; r < nrows; ++r)
            for (int k = 0; k < 3; ++k) mu[k] += uvw[3 * r + k] * freqs[ch] / C;
    for (int k = 0; k < 3; ++k) mu[k] /= (double)ns;
    /* SVD_MFS.py:64-68: PCA(2) on D -> covariance of centred samples. */
    double cov[3][3] = {{0}};
    for (size_t ch = 0; ch < nchan; ++ch) {
        for (size_t r = 0; r < nrows; ++r) {
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = uvw[3 * r + k] * freqs[ch] / C - mu[k];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) cov[a][b] += d[a] * d[b];
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov[a][b] /= (double)(ns - 1);
    double evec[3][3], eval[3];
    jacobi3(cov, evec, eval);
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (eval[order[j]] > eval[order[i]]) { int t = order[i]; order[i] = order[j]; order[j] = t; }
    double V[3][3];
    for (int c = 0; c < 2; ++c) {
        for (int k = 0; k < 3; ++k) V[c][k] = evec[k][order[c]];
        /* svd_flip(u_based_decision=False): largest-|entry| positive, first
         * index on ties (SVDPreprocessKernels.cu:174-193 does the same). */
        int am = 0;
        for (int k = 1; k < 3; ++k) if (fabs(V[c][k]) > fabs(V[c][am])) am = k;
        if (V[c][am] < 0) for (int k = 0; k < 3; ++k) V[c][k] = -V[c][k];
    }
    /* SVD_MFS.py:103 V2 = V0 x V1 */
    V[2][0] = V[0][1] * V[1][2] - V[0][2] * V[1][1];
    V[2][1] = V[0][2] * V[1][0] - V[0][0] * V[1][2];
    V[2][2] = V[0][0] * V[1][1] - V[0][1] * V[1][0];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) vin[3 * a + b] = V[a][b];
    /* SVD_MFS.py:68, 105-110: B = transform(D) + mean.V  (un-centred) */
    const double m0 = mu[0] * V[0][0] + mu[1] * V[0][1] + mu[2] * V[0][2];
    const double m1 = mu[0] * V[1][0] + mu[1] * V[1][1] + mu[2] * V[1][2];
    for (size_t ch = 0; ch < nchan; ++ch) {
        for (size_t r = 0; r < nrows; ++r) {
            const size_t n = r + ch * nrows;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = uvw[3 * r + k] * freqs[ch] / C - mu[k];
            bin[2 * n] = d[0] * V[0][0] + d[1] * V[0][1] + d[2] * V[0][2] + m0;
            bin[2 * n + 1] = d[0] * V[1][0] + d[1] * V[1][1] + d[2] * V[1][2] + m1;
        }
    }
    return 0;
}

void fo_collapse(const double *xx_re, const double *xx_im, const double *yy_re,
                 const double *yy_im, const unsigned char *flag0,
                 const unsigned char *flag3, const double *w0, const double *w3,
                 size_t n, double *out_re, double *out_im)
{
    /* SVD_MFS.py:77-93 */
    for (size_t i = 0; i < n; ++i) {
        const double a = flag0[i] ? 0.0 : w0[i];
        const double b = flag3[i] ? 0.0 : w3[i];
        const double s = a + b;
        if (s > 0.0) {
            out_re[i] = (xx_re[i] * a + yy_re[i] * b) / s;
            out_im[i] = (xx_im[i] * a + yy_im[i] * b) / s;
        } else {
            out_re[i] = 0.0;
            out_im[i] = 0.0;
        }
    }
}

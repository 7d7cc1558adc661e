/*
 * bf_oracle.c -- CPU restatement of the reference bilateral filter.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The product path (opencuda_amd) never
 * links or calls it and fails loudly when its HIP library is missing.
 *
 * What it restates (paths relative to the LitLeo/OpenCUDA tree):
 *   orc_gauss_template   okano_3_0/TemplateFactory.cu:1092-1168 (ring order,
 *                        weights exp((float)(-(i*i+j*j)) / sigma22)),
 *                        sigma22 = 2*sigma*sigma in float (:1104)
 *   orc_euclid_template  okano_3_0/TemplateFactory.cu:1323-1373
 *                        (E[d] = exp((float)(-(d*d)) / sigma22), d = 0..size-1)
 *   orc_filter           okano_3_0/BilateralFilter.cu:71-199 (kernel body),
 *                        launch geometry :323-328, texture clamp :18,:61-64.
 *                        mode ORC_REF reproduces the kernel as written:
 *                        4-row groups read the window of the group's first
 *                        row (:119-120,:137,:154,:171) and run (2r+1)^2+1
 *                        taps (:78,:113).  mode ORC_STD is the textbook
 *                        variant (own window, (2r+1)^2 taps) that the
 *                        surrogate is trained against.
 *   orc_filter_f64       same taps in double precision with the LUT formulas
 *                        evaluated in double and no truncation: the
 *                        real-valued result every u8 output is compared to.
 *
 * Arithmetic contract: built with -ffp-contract=off and without fast-math,
 * so f = g*E, t += f*p and s += f round separately in tap order, exactly as
 * the HIP kernels do.  The u8 conversion truncates toward zero and clamps
 * to [0,255] with NaN -> 0, which is what a saturating float->u8 conversion
 * (CUDA's cvt.rzi.u8.f32) gives for the reference's (unsigned char)(t/sum).
 *
 * Mutations: orc_set_mutation() plants a known mistake (tests only) so the
 * parity checks can show they detect it.  Mutation 0 is the faithful code.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_REF 0
#define ORC_STD 1

enum {
    MUT_NONE = 0,
    MUT_DROP_EXTRA_TAP = 1,   /* REF runs (2r+1)^2 taps instead of +1      */
    MUT_SWAP_DYDX = 2,        /* first template coordinate used as column  */
    MUT_OWN_ROW_WINDOW = 3,   /* REF rows 1..3 read their own window       */
    MUT_ROUND_NEAREST = 4,    /* round instead of truncate                 */
    MUT_FUSED_MULADD = 5,     /* t = fmaf(f, p, t): one rounding, not two  */
    MUT_CENTRE_ROW0 = 6,      /* every row compares against centre[0]      */
    MUT_CLAMP_TRIMMED = 7     /* clamp to the trimmed (output) size        */
};

static int g_mutation = MUT_NONE;

/* Test hook: when set, every written pixel also stores its quotient t/s
 * (before truncation) at g_qout[y * g_qpitch + x]. */
static float *g_qout = NULL;
static size_t g_qpitch = 0;

void orc_set_mutation(int m) { g_mutation = m; }
int orc_get_mutation(void) { return g_mutation; }

int orc_version(void) { return 1; }

/* ------------------------------------------------------------------ LUTs */

/* TemplateFactory.cu:1092-1168.  size is the template edge (odd, >= 3);
 * writes size*size (dy,dx) pairs into offsets and weights into weights.
 * Returns the point count, or -1 on an illegal parameter
 * (TemplateFactory.cu:1057-1076). */
int orc_gauss_template(float sigma, int size, int *offsets, float *weights)
{
    if (size < 3 || (size % 2) == 0 || fabsf(sigma) < 1.0e-8f)
        return -1;
    float sigma22 = 2 * sigma * sigma;            /* (2*sigma)*sigma, float */
    int radius = size / 2;
    int n = 0;
    offsets[2 * n] = 0;
    offsets[2 * n + 1] = 0;
    weights[n++] = 1.0f;
    for (int i = 1; i <= radius; i++) {
        for (int j = -i + 1; j <= i; j++) {
            float w = expf((float)(-(i * i + j * j)) / sigma22);
            const int pts[4][2] = { { i, j }, { -i, j }, { j, i }, { j, -i } };
            for (int k = 0; k < 4; k++) {
                offsets[2 * n] = pts[k][0];
                offsets[2 * n + 1] = pts[k][1];
                weights[n++] = w;
            }
        }
    }
    return n;
}

/* TemplateFactory.cu:1323-1373: E[i] = exp(-(i*i) / sigma22), i < size. */
int orc_euclid_template(float sigma, int size, float *weights)
{
    if (size < 1 || fabsf(sigma) < 1.0e-8f)
        return -1;
    float sigma22 = 2 * sigma * sigma;
    for (int i = 0; i < size; i++)
        weights[i] = expf((float)(-(i * i)) / sigma22);
    return size;
}

/* --------------------------------------------------------------- sampler */

/* Point-sampled, clamp-to-edge fetch of the u8 texture bound to the input
 * ROI sub-image (BilateralFilter.cu:18, :61-64): w,h are the UNTRIMMED
 * input ROI size. */
static inline unsigned char tex2d(const unsigned char *base, size_t pitch,
                                  int w, int h, int x, int y)
{
    x = x < 0 ? 0 : (x >= w ? w - 1 : x);
    y = y < 0 ? 0 : (y >= h ? h - 1 : y);
    return base[(size_t)y * pitch + (size_t)x];
}

static inline unsigned char to_u8(float q)
{
    if (g_mutation == MUT_ROUND_NEAREST)
        q = q + 0.5f;
    if (!(q > 0.0f))
        return 0;                 /* NaN (0/0) and non-positive */
    if (q >= 255.0f)
        return 255;
    return (unsigned char)(int)q; /* truncation toward zero */
}

/* ----------------------------------------------------- the transcription */

/* One thread of _bilateralFilterKer (BilateralFilter.cu:71-199) in REF
 * mode: column dstc, rows dstr..dstr+3, window of row dstr. */
static void ref_thread(const unsigned char *in, size_t inpitch, int inw,
                       int inh, unsigned char *out, size_t outpitch, int W,
                       int H, int radius, const int *tpl, const float *g,
                       const float *E, int dstc, int dstr, int ybase)
{
    int gi = (2 * radius + 1) * (2 * radius + 1);
    if (g_mutation == MUT_DROP_EXTRA_TAP)
        gi -= 1;
    if (dstc >= W || dstr >= H)
        return;
    int cw = inw, ch = inh;
    if (g_mutation == MUT_CLAMP_TRIMMED) {
        cw = W;
        ch = H;
    }
    size_t dstidx = (size_t)(dstr - ybase) * outpitch + (size_t)dstc;
    float sum[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    float t[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    unsigned char center[4];
    for (int k = 0; k < 4; k++)
        center[k] = tex2d(in, inpitch, cw, ch, dstc, dstr + k);
    if (g_mutation == MUT_CENTRE_ROW0)
        center[1] = center[2] = center[3] = center[0];

    for (int col = 0; col <= gi; col++) {
        int i = tpl[2 * col], j = tpl[2 * col + 1];
        if (g_mutation == MUT_SWAP_DYDX) {
            int tmp = i;
            i = j;
            j = tmp;
        }
        for (int k = 0; k < 4; k++) {
            if (k > 0 && dstr + k >= H)
                break;            /* the kernel's `continue` (:133,:150,:167) */
            int row = dstr + i;
            if (g_mutation == MUT_OWN_ROW_WINDOW)
                row += k;
            unsigned char curpix = tex2d(in, inpitch, cw, ch, dstc + j, row);
            unsigned char eu = curpix > center[k] ? curpix - center[k]
                                                  : center[k] - curpix;
            float factor = g[col] * E[eu];
            if (g_mutation == MUT_FUSED_MULADD)
                t[k] = fmaf(factor, (float)curpix, t[k]);
            else
                t[k] += factor * curpix;
            sum[k] += factor;
        }
    }
    for (int k = 0; k < 4; k++) {
        if (dstr + k >= H)
            return;
        float q = t[k] / sum[k];
        if (g_qout)
            g_qout[(size_t)(dstr + k - ybase) * g_qpitch + (size_t)dstc] = q;
        out[dstidx] = to_u8(q);
        dstidx += outpitch;
    }
}

/* STD mode: one output pixel with its own (2r+1)^2 window. */
static void std_pixel(const unsigned char *in, size_t inpitch, int inw,
                      int inh, unsigned char *out, size_t outpitch, int W,
                      int H, int radius, const int *tpl, const float *g,
                      const float *E, int x, int y, int ybase)
{
    (void)W;
    (void)H;
    int gi = (2 * radius + 1) * (2 * radius + 1);
    unsigned char c = tex2d(in, inpitch, inw, inh, x, y);
    float t = 0.0f, s = 0.0f;
    for (int col = 0; col < gi; col++) {
        int i = tpl[2 * col], j = tpl[2 * col + 1];
        if (g_mutation == MUT_SWAP_DYDX) {
            int tmp = i;
            i = j;
            j = tmp;
        }
        unsigned char p = tex2d(in, inpitch, inw, inh, x + j, y + i);
        unsigned char d = p > c ? p - c : c - p;
        float f = g[col] * E[d];
        if (g_mutation == MUT_FUSED_MULADD)
            t = fmaf(f, (float)p, t);
        else
            t += f * p;
        s += f;
    }
    float q = t / s;
    if (g_qout)
        g_qout[(size_t)(y - ybase) * g_qpitch + (size_t)x] = q;
    out[(size_t)(y - ybase) * outpitch + (size_t)x] = to_u8(q);
}

/* Filter rows [ybegin, yend) of the W x H region (ybegin a multiple of 4
 * in REF mode, so groups stay those of the whole region).  in/inpitch/
 * inw/inh describe the input ROI sub-image (untrimmed: the clamp extent);
 * out/outpitch receive row y at out + (y - ybegin) * outpitch; W,H = min
 * of both ROI sizes (BilateralFilter.cu:312-320).  tpl/g: the spatial
 * template (>= (2r+1)^2+1 entries in REF mode); E: 256-entry range LUT.
 * nthreads <= 0 uses the OpenMP default. */
int orc_filter_rows(const unsigned char *in, size_t inpitch, int inw, int inh,
                    unsigned char *out, size_t outpitch, int W, int H, int radius,
                    const int *tpl, const float *g, const float *E, int mode,
                    int nthreads, int ybegin, int yend)
{
    if (!in || !out || !tpl || !g || !E)
        return -2;
    if (radius <= 0 || W <= 0 || H <= 0 || W > inw || H > inh)
        return -1;
    if (ybegin < 0 || yend > H || ybegin >= yend || (mode == ORC_REF && ybegin % 4))
        return -1;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    if (mode == ORC_REF) {
        /* Threads of groups that start inside the band; rows past yend of
         * the last group are not written (the thread's row guard uses
         * Hband = min(H, yend) for writes only via the band check below). */
        long g0 = ybegin / 4, g1 = (yend + 3) / 4;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
        for (long gy = g0; gy < g1; gy++)
            for (int x = 0; x < W; x++)
                ref_thread(in, inpitch, inw, inh, out, outpitch, W,
                           H < yend ? H : yend, radius, tpl, g, E, x, (int)gy * 4,
                           ybegin);
    } else if (mode == ORC_STD) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
        for (int y = ybegin; y < yend; y++)
            for (int x = 0; x < W; x++)
                std_pixel(in, inpitch, inw, inh, out, outpitch, W, H, radius,
                          tpl, g, E, x, y, ybegin);
    } else {
        return -1;
    }
    return 0;
}

/* The whole W x H region. */
int orc_filter(const unsigned char *in, size_t inpitch, int inw, int inh,
               unsigned char *out, size_t outpitch, int W, int H, int radius,
               const int *tpl, const float *g, const float *E, int mode,
               int nthreads)
{
    return orc_filter_rows(in, inpitch, inw, inh, out, outpitch, W, H, radius,
                           tpl, g, E, mode, nthreads, 0, H);
}

/* orc_filter, also storing the float quotient t/s of every output pixel
 * into q (W x H, row pitch qpitch floats).  Single-threaded. */
int orc_filter_q(const unsigned char *in, size_t inpitch, int inw, int inh,
                 unsigned char *out, size_t outpitch, int W, int H, int radius,
                 const int *tpl, const float *g, const float *E, int mode,
                 float *q, size_t qpitch)
{
    g_qout = q;
    g_qpitch = qpitch;
    int rc = orc_filter(in, inpitch, inw, inh, out, outpitch, W, H, radius, tpl,
                        g, E, mode, 1);
    g_qout = NULL;
    return rc;
}

/* ------------------------------------------------ FP64 real-valued oracle */

/* Same taps and mode as orc_filter, weights from the closed forms in
 * double (g = exp(-q / (2 ss^2)), E = exp(-d^2 / (2 sr^2))), result
 * t/s in double with no truncation.  Rows [ybegin, yend) of the W x H
 * output are computed into out (row y at out + (y - ybegin) * W). */
int orc_filter_f64(const unsigned char *in, size_t inpitch, int inw, int inh,
                   double *out, int W, int H, int radius, const int *tpl,
                   double sigma_space, double sigma_range, int mode,
                   int nthreads, int ybegin, int yend)
{
    if (ybegin < 0)
        ybegin = 0;
    if (yend < 0 || yend > H)
        yend = H;
    if (!in || !out || !tpl)
        return -2;
    if (radius <= 0 || W <= 0 || H <= 0 || W > inw || H > inh)
        return -1;
    int gi = (2 * radius + 1) * (2 * radius + 1);
    int ntaps = mode == ORC_REF ? gi + 1 : gi;
    double *gw = (double *)malloc(sizeof(double) * (size_t)ntaps);
    double Ed[256];
    if (!gw)
        return -6;
    for (int c = 0; c < ntaps; c++) {
        double q = (double)(tpl[2 * c] * tpl[2 * c] +
                            tpl[2 * c + 1] * tpl[2 * c + 1]);
        gw[c] = exp(-q / (2.0 * sigma_space * sigma_space));
    }
    for (int d = 0; d < 256; d++)
        Ed[d] = exp(-(double)(d * d) / (2.0 * sigma_range * sigma_range));
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int y = ybegin; y < yend; y++) {
        int y0 = mode == ORC_REF ? y - (y % 4) : y;
        for (int x = 0; x < W; x++) {
            int c = tex2d(in, inpitch, inw, inh, x, y);
            double t = 0.0, s = 0.0;
            for (int k = 0; k < ntaps; k++) {
                int p = tex2d(in, inpitch, inw, inh, x + tpl[2 * k + 1],
                              y0 + tpl[2 * k]);
                double f = gw[k] * Ed[p > c ? p - c : c - p];
                t += f * p;
                s += f;
            }
            out[(size_t)(y - ybegin) * (size_t)W + (size_t)x] = s > 0.0 ? t / s : 0.0;
        }
    }
    free(gw);
    return 0;
}

/* OpenMP's default thread count on this host (1 without OpenMP): recorded
 * next to the CPU baseline's timings. */
int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

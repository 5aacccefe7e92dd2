/*
 * tdc_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's time-domain-correlation (TDC) image
 * formation path, used as the checker for the HIP library. Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load this
 * file's library. The product path never calls it.
 *
 * Two families of functions live here:
 *
 *   orc_image()        FP64 direct sum of the reference formula (the oracle).
 *   orc_refmethod()    FP32 restatement of the reference *method*: per pixel,
 *                      per 128-chirp chunk, materialise the window array, the
 *                      reference array and the product array, then sum. This
 *                      is the CPU baseline ("port") and a behaviour check.
 *
 * Reference files followed (paths relative to the reference root):
 *   tdc-lib/src/tdckernels.cu      :32-50 (range window), :98-130 (window),
 *                                  :196-216 (reference response),
 *                                  :260-273 (conjugate multiply),
 *                                  :309-314 (sum + add to pixel)
 *   tdc-lib/include/windowing.h    :19-58  (Doppler centroid/frequency/window)
 *   tdc-lib/include/gpumath.h      :17-87  (norm, dot, quaternion helpers)
 *   tdc-lib/include/tuning.h       :18-38  (chunk size 128, A = 25/46)
 *   tdc-lib/src/tdcprocessor.cpp   :59-167 (driver loop), :204-208 (centre
 *                                  wavelength), :320-383 (staging + 9-probe
 *                                  skip test)
 *
 * Parity pins (what checks this file, in tests/test_oracle.py):
 *   orc_q_prod / orc_q_rot / orc_norm3 / orc_q_conj: the reference's own
 *       known answers (tdctest/mathtest.cpp:12-137) plus rotation invariants.
 *   orc_dot3: closed forms (unit-axis projections), since the reference test
 *       asserts a recomputed expression rather than the function's output.
 *   orc_range_window / orc_taylor_window: scipy.signal.windows
 *       (general_hamming, hann, taylor with sym=False).
 *   orc_doppler_*: closed forms (velocity perpendicular / parallel to the line
 *       of sight; window value 1 at the centre, 2A-1 at the band edge, 0 out).
 *   orc_image: point-target closed form (|I(p0)| = |a| * sum of windows),
 *       linearity, chirp-permutation invariance, an mpmath brute force on
 *       tiny scenes, and the skip heuristic on closed-form Doppler geometry.
 *   orc_refmethod: agreement with orc_image to FP32 tolerance on the same
 *       inputs (including skip and Doppler-window cases).
 */
#define _GNU_SOURCE 1 /* sincos() */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_CHUNK 128 /* tuning.h:18 PRI_CHUNKSIZE */
#define ORC_C_EXACT 299792458.0
#define ORC_C_F32 299792458.0F /* tdckernels.cu:27, rounds to 299792448 */
#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* gpumath.h helpers, FP64                                                   */
/* ------------------------------------------------------------------------ */

/* gpumath.h:17-25 */
double orc_norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

/* gpumath.h:46-49 */
double orc_dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* gpumath.h:54-61: scalar part first, vector part negated */
void orc_q_conj(const double *q, double *out)
{
    out[0] = q[0];
    out[1] = -q[1];
    out[2] = -q[2];
    out[3] = -q[3];
}

/* gpumath.h:66-72: Hamilton product, scalar part in element 0 */
void orc_q_prod(const double *a, const double *b, double *out)
{
    double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    double r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    double r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
    out[3] = r3;
}

/* gpumath.h:77-87: v' = vector part of q (0,v) q* */
void orc_q_rot(const double *q, const double *v, double *out)
{
    double pv[4] = {0.0, v[0], v[1], v[2]};
    double qc[4], t[4], r[4];
    orc_q_conj(q, qc);
    orc_q_prod(q, pv, t);
    orc_q_prod(t, qc, r);
    out[0] = r[1];
    out[1] = r[2];
    out[2] = r[3];
}

/* windowing.h:19-24 */
double orc_doppler_centroid(const double *vel, const double *ant, double lambda)
{
    return 2.0 / lambda * orc_dot3(vel, ant);
}

/* windowing.h:29-37: uses the radar-to-target vector target - pos */
double orc_doppler_freq(const double *pos, const double *vel, const double *target,
                        double lambda)
{
    double d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    return 2.0 / lambda * orc_dot3(vel, d) / orc_norm3(d[0], d[1], d[2]);
}

/* windowing.h:43-58. A NaN difference fails the band test and gives 0. */
double orc_doppler_window(double fd, double fc, double bw, int taper, double a)
{
    double delta = fd - fc;
    if (fabs(delta) <= bw / 2.0) {
        if (taper) {
            return a - (1.0 - a) * cos(2.0 * ORC_PI * delta / bw - ORC_PI);
        }
        return 1.0;
    }
    return 0.0;
}

/* tdckernels.cu:32-50: periodic generalised Hamming window, denominator Ns */
void orc_range_window(int ns, double a, int apply, double *out)
{
    for (int m = 0; m < ns; ++m) {
        out[m] = apply ? a - (1.0 - a) * cos(2.0 * ORC_PI * (double) m / (double) ns) : 1.0;
    }
}

/*
 * Taylor window (T. T. Taylor, 1955; the form used for SAR apodisation, e.g.
 * Carrara, Goodman and Majewski, "Spotlight Synthetic Aperture Radar", 1995):
 *   w(n) = 1 + 2 * sum_{m=1}^{nbar-1} F_m cos(2 pi m (n - N/2 + 1/2) / N),
 * with A = acosh(10^(sll/20)) / pi, sigma^2 = nbar^2 / (A^2 + (nbar - 1/2)^2),
 *   F_m = (-1)^(m+1) prod_{i=1}^{nbar-1} (1 - m^2 / (sigma^2 (A^2 + (i-1/2)^2)))
 *         / (2 prod_{i != m} (1 - m^2 / i^2)),
 * normalised to 1 at the centre. periodic != 0 gives the DFT-even variant
 * (an (ns+1)-point symmetric window with its last point dropped). No
 * counterpart in the reference (it has only the compile-time Hamming window);
 * pinned against scipy.signal.windows.taylor.
 */
void orc_taylor_window(int ns, int nbar, double sll, int periodic, double *out)
{
    int n_full = periodic ? ns + 1 : ns;
    double b = pow(10.0, sll / 20.0);
    double a = acosh(b) / ORC_PI;
    double s2 = (double) (nbar * nbar) / (a * a + (nbar - 0.5) * (nbar - 0.5));
    double fm[64];
    if (nbar > 64) nbar = 64;
    for (int m = 1; m < nbar; ++m) {
        double num = (m % 2 == 1) ? 1.0 : -1.0;
        double den = 2.0;
        for (int i = 1; i < nbar; ++i) {
            num *= 1.0 - (double) (m * m) / s2 / (a * a + (i - 0.5) * (i - 0.5));
            if (i != m) den *= 1.0 - (double) (m * m) / (double) (i * i);
        }
        fm[m - 1] = num / den;
    }
    double centre = 0.0;
    {
        double x = (n_full - 1) / 2.0;
        double s = 1.0;
        for (int m = 1; m < nbar; ++m)
            s += 2.0 * fm[m - 1] * cos(2.0 * ORC_PI * m * (x - n_full / 2.0 + 0.5) / n_full);
        centre = s;
    }
    for (int n = 0; n < ns; ++n) {
        double s = 1.0;
        for (int m = 1; m < nbar; ++m)
            s += 2.0 * fm[m - 1] * cos(2.0 * ORC_PI * m * (n - n_full / 2.0 + 0.5) / n_full);
        out[n] = s / centre;
    }
}

/* tdcprocessor.cpp:204-208: lambda_c = c / ((f0 + f_end) / 2) */
double orc_center_wavelength(double t_first, double t_last, double mod_rate, double start_freq)
{
    double end_freq = start_freq + mod_rate * (t_last - t_first);
    return ORC_C_EXACT / ((start_freq + end_freq) / 2.0);
}

/* ------------------------------------------------------------------------ */
/* Scene description shared by the FP64 oracle                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int npri;
    int ns;
    const double *raw;          /* [npri][ns][2] complex (re, im) */
    const double *sample_times; /* [ns] seconds within the chirp */
    const double *pos;          /* [npri][ns][3] or [npri][3] (pos_per_sample = 0) */
    int pos_per_sample;
    const double *vel;          /* [npri][ns][3]; only read when doppler != 0 */
    const double *att;          /* [npri][ns][4]; only read when doppler != 0 */
    double mod_rate;            /* K, Hz/s */
    double start_freq;          /* f0, Hz */
    double lambda_c;            /* centre wavelength used by the Doppler terms */
    double bore[3];             /* body boresight */
    const double *range_win;    /* [ns] */
    int doppler;                /* 0: no Doppler window and no skip test */
    double dop_center;
    double dop_bw;
    int centroid;
    int taper;
    double az_a;                /* azimuth window parameter, 25/46 */
    double c;                   /* propagation speed used in the phase */
} orc_scene;

static const double *pos_at(const orc_scene *s, int n, int m)
{
    return s->pos_per_sample ? s->pos + ((size_t) n * s->ns + m) * 3 : s->pos + (size_t) n * 3;
}

/* Doppler centre for one (chirp, sample): windowing.h:19-24 or the fixed one. */
static double fc_at(const orc_scene *s, const double *vel, const double *att)
{
    if (!s->centroid) return s->dop_center;
    double ant[3];
    orc_q_rot(att, s->bore, ant);
    return orc_doppler_centroid(vel, ant, s->lambda_c);
}

/*
 * Skip test for one (pixel, chunk), tdcprocessor.cpp:347-383. Probes rows
 * max(i*64-1, 0) of the 128-row staging buffer and samples
 * max(j*(ns/2)-1, 0), i, j in {0,1,2}. The staging buffer keeps the rows of
 * the previous chunk beyond the valid rows of a partial last chunk
 * (tdcprocessor.cpp:325-345 copies only the valid rows), so such probes read
 * the previous chunk's row; a partial first chunk reads freshly allocated
 * page-locked memory, taken here as zeros. The minimum is taken the way
 * std::min_element does it: a NaN first probe makes the result NaN, a later
 * NaN is never selected. Returns 1 when the chunk is processed.
 */
int orc_chunk_processed(const orc_scene *s, int chunk, const double *target)
{
    double mags[9];
    int pri_step = ORC_CHUNK / 2;
    int sample_step = s->ns / 2;
    int valid = s->npri - chunk * ORC_CHUNK;
    if (valid > ORC_CHUNK) valid = ORC_CHUNK;
    for (int i = 0; i < 3; ++i) {
        int row = i * pri_step - 1;
        if (row < 0) row = 0;
        for (int j = 0; j < 3; ++j) {
            int smp = j * sample_step - 1;
            if (smp < 0) smp = 0;
            double pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, att[4] = {0, 0, 0, 0};
            int pri = -1;
            if (row < valid) pri = chunk * ORC_CHUNK + row;
            else if (chunk > 0) pri = (chunk - 1) * ORC_CHUNK + row;
            if (pri >= 0) {
                size_t e = (size_t) pri * s->ns + smp;
                memcpy(pos, pos_at(s, pri, smp), sizeof pos);
                memcpy(vel, s->vel + e * 3, sizeof vel);
                memcpy(att, s->att + e * 4, sizeof att);
            }
            double fc = fc_at(s, vel, att);
            mags[i * 3 + j] = fabs(orc_doppler_freq(pos, vel, target, s->lambda_c) - fc);
        }
    }
    double mn = mags[0];
    for (int k = 1; k < 9; ++k)
        if (mags[k] < mn) mn = mags[k];
    return mn <= s->dop_bw / 2.0;
}

/*
 * FP64 image for a list of pixels. out[2*i], out[2*i+1] = I(grid[i]).
 *   I(p) = sum over processed chunks c, chirps n in c, samples m of
 *          raw[n,m] * wr[m] * wa[n,m](p) * exp(+j 2 pi * 2 (f0 + K t_m) |pos[n,m] - p| / c)
 * tdckernels.cu:206-215 builds ref = w exp(-j 4 pi f R / c) and :270-272
 * multiplies raw by conj(ref), which is the + sign above.
 */
void orc_image(const orc_scene *s, const double *grid, int npx, double *out, int nthreads)
{
    int nchunks = (s->npri + ORC_CHUNK - 1) / ORC_CHUNK;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int p = 0; p < npx; ++p) {
        const double *tg = grid + (size_t) p * 3;
        double acc_re = 0.0, acc_im = 0.0;
        for (int c = 0; c < nchunks; ++c) {
            if (s->doppler && !orc_chunk_processed(s, c, tg)) continue;
            int n_end = (c + 1) * ORC_CHUNK;
            if (n_end > s->npri) n_end = s->npri;
            double ch_re = 0.0, ch_im = 0.0;
            for (int n = c * ORC_CHUNK; n < n_end; ++n) {
                for (int m = 0; m < s->ns; ++m) {
                    size_t e = (size_t) n * s->ns + m;
                    const double *ps = pos_at(s, n, m);
                    double dx = ps[0] - tg[0], dy = ps[1] - tg[1], dz = ps[2] - tg[2];
                    double r = orc_norm3(dx, dy, dz);
                    double w = s->range_win[m];
                    if (s->doppler) {
                        const double *v = s->vel + e * 3;
                        double fc = fc_at(s, v, s->att + e * 4);
                        double fd = orc_doppler_freq(ps, v, tg, s->lambda_c);
                        w *= orc_doppler_window(fd, fc, s->dop_bw, s->taper, s->az_a);
                    }
                    if (w == 0.0) continue;
                    double f = s->start_freq + s->mod_rate * s->sample_times[m];
                    double cyc = 2.0 * f * r / s->c;
                    cyc -= floor(cyc);
                    double sn, cs;
                    sincos(2.0 * ORC_PI * cyc, &sn, &cs);
                    double rr = s->raw[e * 2], ri = s->raw[e * 2 + 1];
                    ch_re += w * (rr * cs - ri * sn);
                    ch_im += w * (rr * sn + ri * cs);
                }
            }
            acc_re += ch_re;
            acc_im += ch_im;
        }
        out[(size_t) p * 2] = acc_re;
        out[(size_t) p * 2 + 1] = acc_im;
    }
}

/* ------------------------------------------------------------------------ */
/* FP32 restatement of the reference method (CPU baseline "port")            */
/* ------------------------------------------------------------------------ */

typedef struct {
    float x, y, z;
} f3;

static float f_norm3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

static void f_qprod(const float *a, const float *b, float *out)
{
    /* gpumath.h:66-72 in FP32 */
    float r0 = (a[0] * b[0]) - (a[1] * b[1]) - (a[2] * b[2]) - (a[3] * b[3]);
    float r1 = (a[0] * b[1]) + (a[1] * b[0]) + (a[2] * b[3]) - (a[3] * b[2]);
    float r2 = (a[0] * b[2]) - (a[1] * b[3]) + (a[2] * b[0]) + (a[3] * b[1]);
    float r3 = (a[0] * b[3]) + (a[1] * b[2]) - (a[2] * b[1]) + (a[3] * b[0]);
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
    out[3] = r3;
}

static f3 f_qrot(const float *q, const float *v)
{
    /* gpumath.h:77-87 in FP32: vector part of q (0,v) q* */
    float pv[4] = {0.0F, v[0], v[1], v[2]};
    float qc[4] = {q[0], -q[1], -q[2], -q[3]};
    float t[4], r[4];
    f_qprod(q, pv, t);
    f_qprod(t, qc, r);
    f3 o = {r[1], r[2], r[3]};
    return o;
}

static float f_dop_freq(const float *pos, const float *vel, const float *tg, float lambda)
{
    float d0 = tg[0] - pos[0], d1 = tg[1] - pos[1], d2 = tg[2] - pos[2];
    return 2.0F / lambda * (vel[0] * d0 + vel[1] * d1 + vel[2] * d2) / f_norm3(d0, d1, d2);
}

static float f_dop_window(float fd, float fc, float bw, int taper)
{
    const float a = (float) (25.0 / 46.0);
    const float pi_f = 3.141592653589F;
    float delta = fd - fc;
    if (fabs(delta) <= bw / 2.0) {
        if (taper) return a - ((1.0F - a) * cosf((2.0F * pi_f * delta / bw) - pi_f));
        return 1.0F;
    }
    return 0.0F;
}

/* sin and cos of pi*x in FP32 with exact reduction of x modulo 2 */
static void f_sincospi(float x, float *s, float *c)
{
    float r = fmodf(x, 2.0F);
    double a = ORC_PI * (double) r;
    *s = (float) sin(a);
    *c = (float) cos(a);
}

/* pairwise sum of n complex values (re, im interleaved) */
static void pairwise_sum(const float *v, size_t n, float *re, float *im)
{
    if (n <= 16) {
        float sr = 0.0F, si = 0.0F;
        for (size_t i = 0; i < n; ++i) {
            sr += v[2 * i];
            si += v[2 * i + 1];
        }
        *re = sr;
        *im = si;
        return;
    }
    size_t h = n / 2;
    float ar, ai, br, bi;
    pairwise_sum(v, h, &ar, &ai);
    pairwise_sum(v + 2 * h, n - h, &br, &bi);
    *re = ar + br;
    *im = ai + bi;
}

/*
 * Reference method in FP32, for a list of pixels. Inputs use the reference's
 * own float layouts (tdcprocessor.h:57-107). For each pixel and chunk:
 *   skip test (FP32, tdcprocessor.cpp:347-383, same staging quirk as above),
 *   window[n,m]  = wr[m] * dopplerWindow(...)       (tdckernels.cu:98-130)
 *   ref[n,m]     = window * (cospi phi, sinpi phi),  phi = -4 R (f / c_f)
 *                                                    (tdckernels.cu:196-216)
 *   prod[n,m]    = raw * conj(ref)                   (tdckernels.cu:260-273)
 *   pixel       += sum(prod)                         (tdckernels.cu:309-314)
 * wr is passed in (the reference's periodic Hamming, or ones).
 */
void orc_refmethod(const float *raw, const float *sample_times, const float *pos, const float *vel,
                   const float *att, int npri, int ns, float mod_rate, float start_freq,
                   const float *bore, float lambda_c, const float *range_win, const float *grid,
                   int npx, float dop_center, float dop_bw, int centroid, int taper,
                   int use_doppler, float *out, int nthreads)
{
    int nchunks = (npri + ORC_CHUNK - 1) / ORC_CHUNK;
    size_t chunk_elems = (size_t) ORC_CHUNK * ns;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        float *window = (float *) malloc(chunk_elems * sizeof(float));
        float *ref = (float *) malloc(chunk_elems * 2 * sizeof(float));
#pragma omp for schedule(dynamic, 1)
        for (int p = 0; p < npx; ++p) {
            const float *tg = grid + (size_t) p * 3;
            float img_re = 0.0F, img_im = 0.0F;
            for (int c = 0; c < nchunks; ++c) {
                int valid = npri - c * ORC_CHUNK;
                if (valid > ORC_CHUNK) valid = ORC_CHUNK;
                if (use_doppler) {
                    float mags[9];
                    int sample_step = ns / 2;
                    for (int i = 0; i < 3; ++i) {
                        int row = i * (ORC_CHUNK / 2) - 1;
                        if (row < 0) row = 0;
                        for (int j = 0; j < 3; ++j) {
                            int smp = j * sample_step - 1;
                            if (smp < 0) smp = 0;
                            float zp[3] = {0, 0, 0}, zv[3] = {0, 0, 0}, za[4] = {0, 0, 0, 0};
                            const float *pp = zp, *vv = zv, *aa = za;
                            int pri = -1;
                            if (row < valid) pri = c * ORC_CHUNK + row;
                            else if (c > 0) pri = (c - 1) * ORC_CHUNK + row;
                            if (pri >= 0) {
                                size_t e = (size_t) pri * ns + smp;
                                pp = pos + e * 3;
                                vv = vel + e * 3;
                                aa = att + e * 4;
                            }
                            float fc = dop_center;
                            if (centroid) {
                                f3 ant = f_qrot(aa, bore);
                                fc = 2.0F / lambda_c * (vv[0] * ant.x + vv[1] * ant.y + vv[2] * ant.z);
                            }
                            mags[i * 3 + j] = fabsf(f_dop_freq(pp, vv, tg, lambda_c) - fc);
                        }
                    }
                    float mn = mags[0];
                    for (int k = 1; k < 9; ++k)
                        if (mags[k] < mn) mn = mags[k];
                    if (!(mn <= dop_bw / 2.0)) continue;
                }
                /* window array */
                for (int n = 0; n < valid; ++n) {
                    for (int m = 0; m < ns; ++m) {
                        size_t e = (size_t) (c * ORC_CHUNK + n) * ns + m;
                        float az = 1.0F;
                        if (use_doppler) {
                            float fc = dop_center;
                            if (centroid) {
                                f3 ant = f_qrot(att + e * 4, bore);
                                const float *vv = vel + e * 3;
                                fc = 2.0F / lambda_c * (vv[0] * ant.x + vv[1] * ant.y + vv[2] * ant.z);
                            }
                            az = f_dop_window(f_dop_freq(pos + e * 3, vel + e * 3, tg, lambda_c), fc,
                                              dop_bw, taper);
                        }
                        window[(size_t) n * ns + m] = range_win[m] * az;
                    }
                }
                /* reference array */
                for (int n = 0; n < valid; ++n) {
                    for (int m = 0; m < ns; ++m) {
                        size_t e = (size_t) (c * ORC_CHUNK + n) * ns + m;
                        size_t l = (size_t) n * ns + m;
                        const float *ps = pos + e * 3;
                        float r = f_norm3(ps[0] - tg[0], ps[1] - tg[1], ps[2] - tg[2]);
                        float f = fmaf(sample_times[m], mod_rate, start_freq);
                        float phi = -4.0F * r * (f / ORC_C_F32);
                        float sn, cs;
                        f_sincospi(phi, &sn, &cs);
                        ref[2 * l] = cs * window[l];
                        ref[2 * l + 1] = sn * window[l];
                    }
                }
                /* product raw * conj(ref), in place */
                for (int n = 0; n < valid; ++n) {
                    for (int m = 0; m < ns; ++m) {
                        size_t e = (size_t) (c * ORC_CHUNK + n) * ns + m;
                        size_t l = (size_t) n * ns + m;
                        float v1r = raw[2 * e], v1i = raw[2 * e + 1];
                        float v2r = ref[2 * l], v2i = -ref[2 * l + 1];
                        ref[2 * l] = v1r * v2r - v1i * v2i;
                        ref[2 * l + 1] = v1r * v2i + v1i * v2r;
                    }
                }
                float sr, si;
                pairwise_sum(ref, (size_t) valid * ns, &sr, &si);
                img_re += sr;
                img_im += si;
            }
            out[(size_t) p * 2] = img_re;
            out[(size_t) p * 2 + 1] = img_im;
        }
        free(window);
        free(ref);
    }
}

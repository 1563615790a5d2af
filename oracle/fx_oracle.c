/*
 * fx_oracle.c -- CPU oracle for the FX correlator hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see fx_oracle.h).  Built with -ffp-contract=off so
 * that every product and sum is rounded separately, as the reference's
 * one-primitive-per-call IPP sequence is.
 *
 * Citations are to the reference sources (twiG56/difx_tests).
 */
#include "fx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* IPP_PI as the reference's TWO_PI macro uses it (vectordefs.h:8). */
static const double OR_PI = 3.14159265358979323846;
#define OR_TINY 0.00000000001 /* vectordefs.h:9 */

static int ilog2_exact(int n)
{
    int lg = 0;
    if (n <= 0) return -1;
    while ((1 << lg) < n) lg++;
    return ((1 << lg) == n) ? lg : -1;
}

int or_plan_init(or_plan *p, int nchan, double lo, double bw)
{
    int lg = ilog2_exact(nchan);
    if (lg < 2) return -1;
    if (!(lo >= 0.0 && lo < 2147483647.0)) return -2; /* (int)lo, kernel.c:39 */
    p->nchan = nchan;
    p->F = 2 * nchan;                      /* kernel.c:13 */
    p->Ts = 1.0 / (2.0 * bw);              /* kernel.c:14 */
    p->stride = 1 << (lg / 2);             /* kernel.c:31 for square nchan */
    p->sub = 2 * p->stride;                /* kernel.c:36 (real data) */
    p->nstep = p->F / p->sub;
    p->nfstep = nchan / p->stride;
    p->lo = lo;
    p->bw = bw;
    p->fracLO = (lo - (double)(int)lo) > OR_TINY; /* kernel.c:39 */
    p->iscomplex = 0;
    p->origin = 0;
    return 0;
}

int or_plan_init_complex(or_plan *p, int nchan, double lo, double bw)
{
    int rc = or_plan_init(p, nchan, lo, bw);
    if (rc) return rc;
    p->F = nchan;                          /* cfactor 1 (allocDataHost.c:12-19) */
    p->Ts = 1.0 / bw;                      /* no halving for complex sampling */
    p->nstep = p->F / p->sub;
    p->iscomplex = 1;
    return 0;
}

long long or_subint_bytes_complex(int nchan, int numffts)
{
    return (long long)nchan * (numffts + 1) * 4;
}

void or_unpack8complex(const uint8_t *data, int offset, int F, or_cf32 *x0, or_cf32 *x1)
{
    const int8_t *d = (const int8_t *)data + 4 * (size_t)offset;
    for (int i = 0; i < F; i++) {
        x0[i].re = (float)d[4 * i];
        x0[i].im = (float)d[4 * i + 1];
        x1[i].re = (float)d[4 * i + 2];
        x1[i].im = (float)d[4 * i + 3];
    }
}

long long or_subint_bytes(int nchan, int numffts)
{
    /* allocDataHost.c:21 with cfactor=2 (real), nbit=2, nPol=2 */
    return (long long)nchan * 2 * (numffts + 1) * 2 / 8 * 2;
}

/* kernel.c:373-386 */
void or_station_delay(const double c[3], int fftindex, double *mean, double *a, double *b)
{
    double d0, d1, d2;
    d0 = c[0] * fftindex * fftindex + c[1] * fftindex + c[2];
    d1 = c[0] * (fftindex + 0.5) * (fftindex + 0.5) + c[1] * (fftindex + 0.5) + c[2];
    d2 = c[0] * (fftindex + 1.0) * (fftindex + 1.0) + c[1] * (fftindex + 1.0) + c[2];
    *a = d2 - d0;
    *b = d0 + (d1 - (*a * 0.5 + d0)) / 3.0;
    *mean = *a * 0.5 + *b;
}

/* kernel.c:736-756 */
/* kernel.c:736-756 with the sample offset taken relative to a buffer that
 * starts at FFT `origin` (0 for the reference). */
int or_offset_origin(double mean, double fileoffset, double Ts, int g, int origin, int F, int maxoffset,
                     int *offset, double *fracdelay)
{
    double x = (mean - fileoffset) / Ts;
    int n = (int)(x + 0.5);               /* truncation toward zero (Q3) */
    double fd = -(x - n) * Ts;
    int o = (g - origin) * F - n;
    if (o == -1) {
        ++o;
        fd += Ts;
    }
    if (o == maxoffset + 1) {
        --o;
        fd -= Ts;
    }
    *offset = o;
    *fracdelay = fd;
    return !(o < 0 || o > maxoffset);
}

int or_offset(double mean, double fileoffset, double Ts, int g, int F, int maxoffset,
              int *offset, double *fracdelay)
{
    return or_offset_origin(mean, fileoffset, Ts, g, 0, F, maxoffset, offset, fracdelay);
}

/* kernel.c:191-192 */
float or_level(int code)
{
    static const float HiMag = (float)3.3359;
    const float lut4level[4] = {-HiMag, (float)-1.0, (float)1.0, HiMag};
    return lut4level[code & 3];
}

/* kernel.c:397-426: sample q of the stream is in byte q/2, time slot q%2;
 * pol p of slot t is bits 2*(2t+p) (kernel.c:195-200). */
void or_unpack2bit(const uint8_t *data, int offset, int F, float *pol0, float *pol1)
{
    for (int i = 0; i < F; i++) {
        int q = offset + i;
        uint8_t byte = data[q >> 1];
        int t = q & 1;
        pol0[i] = or_level(byte >> (2 * (2 * t + 0)));
        pol1[i] = or_level(byte >> (2 * (2 * t + 1)));
    }
}

void or_sincos_f32(float x, float *s, float *c)
{
    *s = (float)sin((double)x);
    *c = (float)cos((double)x);
}

static inline or_cf32 cmul32(or_cf32 x, or_cf32 y)
{
    or_cf32 r;
    r.re = x.re * y.re - x.im * y.im;
    r.im = x.re * y.im + x.im * y.re;
    return r;
}

static inline or_cf64 cmul64(or_cf64 x, or_cf64 y)
{
    or_cf64 r;
    r.re = x.re * y.re - x.im * y.im;
    r.im = x.re * y.im + x.im * y.re;
    return r;
}

/* Fringe rotator: kernel.c:138-147 (offset tables), kernel.c:428-488. */
void or_fringe_rotator_f32(const or_plan *p, double a, double b, or_cf32 *rot)
{
    const int F = p->F, sub = p->sub, stride = p->stride, nstep = p->nstep;
    const double lo = p->lo;
    int integerdelay = (int)b;
    b -= integerdelay;
    double lofrac = (lo - (int)(lo)) * (double)(integerdelay);
    or_cf32 stepc;
    for (int s = 0; s < sub; s++) {
        double subxoff = ((double)s) / ((double)F);
        double subxval = subxoff * a;
        subxval += b;
        double subphase = subxval * lo;
        if (p->fracLO) subphase += lofrac;
        float arg = (float)((-2.0 * OR_PI) * (subphase - (int)(subphase)));
        or_sincos_f32(arg, &rot[s].im, &rot[s].re);
    }
    for (int k = 1; k < nstep; k++) {
        double stepxoff = ((double)k * (double)stride) / ((double)F); /* Q1 */
        double stepxval = stepxoff * a;
        double stepphase = stepxval * lo;
        float arg = (float)((-2.0 * OR_PI) * (stepphase - (int)(stepphase)));
        or_sincos_f32(arg, &stepc.im, &stepc.re);
        for (int s = 0; s < sub; s++) rot[k * sub + s] = cmul32(rot[s], stepc);
    }
}

void or_fringe_rotator_f64(const or_plan *p, double a, double b, or_cf64 *rot)
{
    const int F = p->F, sub = p->sub, stride = p->stride, nstep = p->nstep;
    const double lo = p->lo;
    int integerdelay = (int)b;
    b -= integerdelay;
    double lofrac = (lo - (int)(lo)) * (double)(integerdelay);
    for (int s = 0; s < sub; s++) {
        double subphase = (((double)s) / ((double)F) * a + b) * lo;
        if (p->fracLO) subphase += lofrac;
        double arg = (-2.0 * OR_PI) * (subphase - (int)(subphase));
        rot[s].re = cos(arg);
        rot[s].im = sin(arg);
    }
    for (int k = 1; k < nstep; k++) {
        double stepphase = (((double)k * (double)stride) / ((double)F) * a) * lo;
        double arg = (-2.0 * OR_PI) * (stepphase - (int)(stepphase));
        or_cf64 stepc = {cos(arg), sin(arg)};
        for (int s = 0; s < sub; s++) rot[k * sub + s] = cmul64(rot[s], stepc);
    }
}

/* Fractional-sample rotator: kernel.c:162-166 (channel frequencies),
 * kernel.c:512-546.  The FP64 fractional delay is narrowed to f32 by the
 * Ipp32f parameter of ippsMulC_32f (Q6). */
void or_frac_rotator_f32(const or_plan *p, double fracdelay, or_cf32 *fr)
{
    const int stride = p->stride, nfstep = p->nfstep, nchan = p->nchan;
    const double bw = p->bw;
    const float fd = (float)fracdelay;
    for (int s = 0; s < stride; s++) {
        float subchf = (float)((2.0 * OR_PI * s * bw) / (double)nchan);
        float arg = subchf * fd;
        or_sincos_f32(arg, &fr[s].im, &fr[s].re);
    }
    for (int k = 1; k < nfstep; k++) {
        float stepchf = (float)((2.0 * OR_PI * k * stride * bw) / (double)nchan);
        float arg = stepchf * fd;
        or_cf32 stepc;
        or_sincos_f32(arg, &stepc.im, &stepc.re);
        for (int s = 0; s < stride; s++) fr[k * stride + s] = cmul32(fr[s], stepc);
    }
}

void or_frac_rotator_f64(const or_plan *p, double fracdelay, or_cf64 *fr)
{
    const int stride = p->stride, nfstep = p->nfstep, nchan = p->nchan;
    const double bw = p->bw;
    for (int s = 0; s < stride; s++) {
        double arg = ((2.0 * OR_PI * s * bw) / (double)nchan) * fracdelay;
        fr[s].re = cos(arg);
        fr[s].im = sin(arg);
    }
    for (int k = 1; k < nfstep; k++) {
        double arg = ((2.0 * OR_PI * k * stride * bw) / (double)nchan) * fracdelay;
        or_cf64 stepc = {cos(arg), sin(arg)};
        for (int s = 0; s < stride; s++) fr[k * stride + s] = cmul64(fr[s], stepc);
    }
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 decimation in time, twiddles from FP64 cos/sin.    */

#define OR_MAX_LG 20
static or_cf32 *tw32[OR_MAX_LG + 1];
static or_cf64 *tw64[OR_MAX_LG + 1];

static void fft_prepare(int lg)
{
    if (tw64[lg]) return;
    int F = 1 << lg;
    or_cf64 *t64 = (or_cf64 *)malloc(sizeof(or_cf64) * (size_t)(F / 2 + 1));
    or_cf32 *t32 = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)(F / 2 + 1));
    for (int k = 0; k <= F / 2; k++) {
        double ang = -2.0 * OR_PI * (double)k / (double)F;
        t64[k].re = cos(ang);
        t64[k].im = sin(ang);
        t32[k].re = (float)t64[k].re;
        t32[k].im = (float)t64[k].im;
    }
    tw32[lg] = t32;
    tw64[lg] = t64;
}

static void fft_prepare_all(int F)
{
    int lg = ilog2_exact(F);
    if (lg >= 0 && lg <= OR_MAX_LG) fft_prepare(lg);
}

static inline unsigned bitrev(unsigned x, int lg)
{
    unsigned r = 0;
    for (int i = 0; i < lg; i++) {
        r = (r << 1) | (x & 1u);
        x >>= 1;
    }
    return r;
}

void or_fft_f32(int F, const or_cf32 *in, or_cf32 *out)
{
    int lg = ilog2_exact(F);
    fft_prepare(lg);
    const or_cf32 *tw = tw32[lg];
    for (int i = 0; i < F; i++) out[bitrev((unsigned)i, lg)] = in[i];
    for (int len = 2; len <= F; len <<= 1) {
        int half = len >> 1, tstep = F / len;
        for (int i = 0; i < F; i += len) {
            for (int j = 0; j < half; j++) {
                or_cf32 w = tw[j * tstep];
                or_cf32 u = out[i + j];
                or_cf32 v = cmul32(out[i + j + half], w);
                out[i + j].re = u.re + v.re;
                out[i + j].im = u.im + v.im;
                out[i + j + half].re = u.re - v.re;
                out[i + j + half].im = u.im - v.im;
            }
        }
    }
}

void or_fft_f64(int F, const or_cf64 *in, or_cf64 *out)
{
    int lg = ilog2_exact(F);
    fft_prepare(lg);
    const or_cf64 *tw = tw64[lg];
    for (int i = 0; i < F; i++) out[bitrev((unsigned)i, lg)] = in[i];
    for (int len = 2; len <= F; len <<= 1) {
        int half = len >> 1, tstep = F / len;
        for (int i = 0; i < F; i += len) {
            for (int j = 0; j < half; j++) {
                or_cf64 w = tw[j * tstep];
                or_cf64 u = out[i + j];
                or_cf64 v = cmul64(out[i + j + half], w);
                out[i + j].re = u.re + v.re;
                out[i + j].im = u.im + v.im;
                out[i + j + half].re = u.re - v.re;
                out[i + j + half].im = u.im - v.im;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* F-stage for one (antenna, FFT): kernel.c:733-788, FP32.                   */

typedef struct {
    float *l0, *l1;         /* unpacked levels, F each */
    or_cf32 *x0, *x1;       /* unpacked complex samples, F each */
    or_cf32 *rot, *u, *y;   /* rotator, rotated input, FFT output (F each) */
    or_cf32 *fr;            /* frac rotator (nchan) */
} f32_scratch;

static int scratch_alloc(f32_scratch *s, int F, int nchan)
{
    s->l0 = (float *)malloc(sizeof(float) * (size_t)F);
    s->l1 = (float *)malloc(sizeof(float) * (size_t)F);
    s->x0 = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)F);
    s->x1 = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)F);
    s->rot = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)F);
    s->u = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)F);
    s->y = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)F);
    s->fr = (or_cf32 *)malloc(sizeof(or_cf32) * (size_t)nchan);
    return (s->l0 && s->l1 && s->x0 && s->x1 && s->rot && s->u && s->y && s->fr) ? 0 : -1;
}

static void scratch_free(f32_scratch *s)
{
    free(s->l0); free(s->l1); free(s->x0); free(s->x1); free(s->rot); free(s->u); free(s->y); free(s->fr);
}

/* Computes the two frac-corrected spectra (channels < nchan) of antenna j at
 * global FFT g into out[2][nchan]; returns validity. */
static int fstage_one_f32(const or_plan *p, const double *c, double fileoffset, int g, int maxoffset,
                          const uint8_t *data, f32_scratch *s, or_cf32 *out,
                          int *offset_out, double *fd_out)
{
    const int F = p->F, nchan = p->nchan;
    double mean, a, b, fd;
    int offset;
    or_station_delay(c, g, &mean, &a, &b);
    int valid = or_offset_origin(mean, fileoffset, p->Ts, g, p->origin, F, maxoffset, &offset, &fd);
    if (offset_out) *offset_out = offset;
    if (fd_out) *fd_out = fd;
    if (!valid) return 0;
    or_fringe_rotator_f32(p, a, b, s->rot);
    or_frac_rotator_f32(p, fd, s->fr);
    if (p->iscomplex)
        or_unpack8complex(data, offset, F, s->x0, s->x1);
    else
        or_unpack2bit(data, offset, F, s->l0, s->l1);
    for (int pol = 0; pol < 2; pol++) {
        if (p->iscomplex) {
            const or_cf32 *x = pol ? s->x1 : s->x0;
            for (int i = 0; i < F; i++) s->u[i] = cmul32(x[i], s->rot[i]);
        } else {
            const float *l = pol ? s->l1 : s->l0;
            for (int i = 0; i < F; i++) {
                s->u[i].re = l[i] * s->rot[i].re;
                s->u[i].im = l[i] * s->rot[i].im;
            }
        }
        or_fft_f32(F, s->u, s->y);
        for (int ch = 0; ch < nchan; ch++) out[pol * nchan + ch] = cmul32(s->fr[ch], s->y[ch]);
    }
    return 1;
}

/* X accumulate for one FFT, kernel.c:829-854: V[b][2l+m] += Y_jl * conj(Y_km). */
static void xacc_f32(int nant, int nchan, const or_cf32 *spec /*[nant][2][nchan]*/, const int *valid,
                     float *vis, int *count, float *autos, int *autocount)
{
    int b = 0;
    for (int j = 0; j < nant - 1; j++) {
        if (!valid[j]) {
            b += nant - (j + 1);
            continue;
        }
        for (int k = j + 1; k < nant; k++) {
            if (!valid[k]) {
                b++;
                continue;
            }
            for (int l = 0; l < 2; l++) {
                for (int m = 0; m < 2; m++) {
                    const or_cf32 *x = spec + ((size_t)j * 2 + l) * nchan;
                    const or_cf32 *y = spec + ((size_t)k * 2 + m) * nchan;
                    float *acc = vis + (((size_t)b * 4 + 2 * l + m) * nchan) * 2;
                    for (int ch = 0; ch < nchan; ch++) {
                        or_cf32 yc = {y[ch].re, -y[ch].im};
                        or_cf32 t = cmul32(x[ch], yc);
                        acc[2 * ch] += t.re;
                        acc[2 * ch + 1] += t.im;
                    }
                }
            }
            count[b]++;
            b++;
        }
    }
    if (autos) {
        for (int j = 0; j < nant; j++) {
            if (!valid[j]) continue;
            for (int l = 0; l < 2; l++)
                for (int m = 0; m < 2; m++) {
                    const or_cf32 *x = spec + ((size_t)j * 2 + l) * nchan;
                    const or_cf32 *y = spec + ((size_t)j * 2 + m) * nchan;
                    float *acc = autos + (((size_t)j * 4 + 2 * l + m) * nchan) * 2;
                    for (int ch = 0; ch < nchan; ch++) {
                        or_cf32 yc = {y[ch].re, -y[ch].im};
                        or_cf32 t = cmul32(x[ch], yc);
                        acc[2 * ch] += t.re;
                        acc[2 * ch + 1] += t.im;
                    }
                }
            if (autocount) autocount[j]++;
        }
    }
}

int or_process_f32(const or_plan *p, int nant, int numffts, int split,
                   const double *delays, const double *fileoffsets, const int *iteration,
                   const uint8_t *const *data,
                   float *vis, int *count, float *autos, int *autocount,
                   float *chan, int *valid_out, int *offsets, double *fracdelays)
{
    const int F = p->F, nchan = p->nchan;
    const int maxoffset = (numffts * split - 1) * F; /* kernel.c:730 */
    fft_prepare_all(F);
    f32_scratch s;
    or_cf32 *spec = (or_cf32 *)calloc((size_t)nant * 2 * nchan, sizeof(or_cf32));
    int *valid = (int *)malloc(sizeof(int) * (size_t)nant);
    if (!spec || !valid || scratch_alloc(&s, F, nchan)) return -1;
    for (int k = 0; k < numffts; k++) {
        for (int j = 0; j < nant; j++) {
            int off;
            double fd;
            valid[j] = fstage_one_f32(p, delays + 3 * j, fileoffsets[j], iteration[k], maxoffset,
                                      data[j], &s, spec + (size_t)j * 2 * nchan, &off, &fd);
            if (offsets) offsets[(size_t)k * nant + j] = off;
            if (fracdelays) fracdelays[(size_t)k * nant + j] = fd;
            if (valid_out) valid_out[(size_t)k * nant + j] = valid[j];
            if (chan) {
                float *dst = chan + ((size_t)k * nant + j) * 2 * nchan * 2;
                if (valid[j])
                    memcpy(dst, spec + (size_t)j * 2 * nchan, sizeof(or_cf32) * 2 * (size_t)nchan);
                else
                    memset(dst, 0, sizeof(or_cf32) * 2 * (size_t)nchan);
            }
        }
        xacc_f32(nant, nchan, spec, valid, vis, count, autos, autocount);
    }
    scratch_free(&s);
    free(spec);
    free(valid);
    return 0;
}

int or_fstage_f32(const or_plan *p, int nant, int numffts, int split,
                  const double *delays, const double *fileoffsets, const int *iteration,
                  const uint8_t *const *data, float *chan, int *valid, int nthreads)
{
    const int F = p->F, nchan = p->nchan;
    const int maxoffset = (numffts * split - 1) * F;
    const long long total = (long long)numffts * nant;
    int err = 0;
    fft_prepare_all(F);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
    {
        f32_scratch s;
        if (scratch_alloc(&s, F, nchan)) {
            err = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (long long t = 0; t < total; t++) {
                int k = (int)(t / nant), j = (int)(t % nant);
                or_cf32 *dst = (or_cf32 *)(chan + (size_t)t * 2 * nchan * 2);
                int v = fstage_one_f32(p, delays + 3 * j, fileoffsets[j], iteration[k], maxoffset,
                                       data[j], &s, dst, NULL, NULL);
                if (!v) memset(dst, 0, sizeof(or_cf32) * 2 * (size_t)nchan);
                valid[t] = v;
            }
        }
        scratch_free(&s);
    }
    (void)nthreads;
    return err ? -1 : 0;
}

int or_xstage_f32(int nant, int nchan, int numffts, const float *chan, const int *valid,
                  float *vis, int *count, int nthreads)
{
    const int nbl = nant * (nant - 1) / 2;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
    for (int b = 0; b < nbl; b++) {
        /* invert the j-major baseline index (kernel.c:831-852) */
        int j = 0, rem = b;
        while (rem >= nant - 1 - j) {
            rem -= nant - 1 - j;
            j++;
        }
        int k = j + 1 + rem;
        for (int f = 0; f < numffts; f++) {
            if (!valid[(size_t)f * nant + j] || !valid[(size_t)f * nant + k]) continue;
            for (int l = 0; l < 2; l++)
                for (int m = 0; m < 2; m++) {
                    const or_cf32 *x = (const or_cf32 *)(chan + (((size_t)f * nant + j) * 2 + l) * nchan * 2);
                    const or_cf32 *y = (const or_cf32 *)(chan + (((size_t)f * nant + k) * 2 + m) * nchan * 2);
                    float *acc = vis + (((size_t)b * 4 + 2 * l + m) * nchan) * 2;
                    for (int ch = 0; ch < nchan; ch++) {
                        or_cf32 yc = {y[ch].re, -y[ch].im};
                        or_cf32 t = cmul32(x[ch], yc);
                        acc[2 * ch] += t.re;
                        acc[2 * ch + 1] += t.im;
                    }
                }
            count[b]++;
        }
    }
    (void)nthreads;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 oracle                                                               */

int or_process_f64(const or_plan *p, int nant, int numffts, int split,
                   const double *delays, const double *fileoffsets, const int *iteration,
                   const uint8_t *const *data,
                   double *vis, int *count, double *autos, int *autocount,
                   double *chan, int *valid_out)
{
    /* OpenMP over antennas (F part) and over the first antenna of each
     * baseline row (X part); every accumulator is written by one thread in
     * the same order as the serial loop, so the result does not depend on
     * the thread count. */
    const int F = p->F, nchan = p->nchan;
    const int maxoffset = (numffts * split - 1) * F;
    fft_prepare_all(F);
#ifdef _OPENMP
    const int nt = omp_get_max_threads();
#else
    const int nt = 1;
#endif
    float *l0 = (float *)malloc(sizeof(float) * (size_t)F * nt);
    float *l1 = (float *)malloc(sizeof(float) * (size_t)F * nt);
    or_cf64 *rot = (or_cf64 *)malloc(sizeof(or_cf64) * (size_t)F * nt);
    or_cf64 *u = (or_cf64 *)malloc(sizeof(or_cf64) * (size_t)F * nt);
    or_cf64 *y = (or_cf64 *)malloc(sizeof(or_cf64) * (size_t)F * nt);
    or_cf64 *fr = (or_cf64 *)malloc(sizeof(or_cf64) * (size_t)nchan * nt);
    or_cf64 *spec = (or_cf64 *)calloc((size_t)nant * 2 * nchan, sizeof(or_cf64));
    int *valid = (int *)malloc(sizeof(int) * (size_t)nant);
    if (!l0 || !l1 || !rot || !u || !y || !fr || !spec || !valid) {
        free(l0); free(l1); free(rot); free(u); free(y); free(fr); free(spec); free(valid);
        return -1;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
#ifdef _OPENMP
        const int tn = omp_get_thread_num();
#else
        const int tn = 0;
#endif
        float *tl0 = l0 + (size_t)tn * F, *tl1 = l1 + (size_t)tn * F;
        or_cf64 *trot = rot + (size_t)tn * F, *tu = u + (size_t)tn * F, *ty = y + (size_t)tn * F;
        or_cf64 *tfr = fr + (size_t)tn * nchan;
        for (int k = 0; k < numffts; k++) {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int j = 0; j < nant; j++) {
                double mean, a, b, fd;
                int off;
                or_station_delay(delays + 3 * j, iteration[k], &mean, &a, &b);
                valid[j] = or_offset_origin(mean, fileoffsets[j], p->Ts, iteration[k], p->origin, F, maxoffset,
                                            &off, &fd);
                if (valid_out) valid_out[(size_t)k * nant + j] = valid[j];
                or_cf64 *sj = spec + (size_t)j * 2 * nchan;
                if (valid[j]) {
                    or_fringe_rotator_f64(p, a, b, trot);
                    or_frac_rotator_f64(p, fd, tfr);
                    const int8_t *cd = (const int8_t *)data[j] + 4 * (size_t)off;
                    if (!p->iscomplex) or_unpack2bit(data[j], off, F, tl0, tl1);
                    for (int pol = 0; pol < 2; pol++) {
                        const float *l = pol ? tl1 : tl0;
                        for (int i = 0; i < F; i++) {
                            if (p->iscomplex) {
                                const or_cf64 x = {(double)cd[4 * i + 2 * pol], (double)cd[4 * i + 2 * pol + 1]};
                                tu[i] = cmul64(x, trot[i]);
                            } else {
                                tu[i].re = (double)l[i] * trot[i].re;
                                tu[i].im = (double)l[i] * trot[i].im;
                            }
                        }
                        or_fft_f64(F, tu, ty);
                        for (int ch = 0; ch < nchan; ch++) sj[pol * nchan + ch] = cmul64(tfr[ch], ty[ch]);
                    }
                }
                if (chan) {
                    double *dst = chan + ((size_t)k * nant + j) * 2 * nchan * 2;
                    if (valid[j])
                        memcpy(dst, sj, sizeof(or_cf64) * 2 * (size_t)nchan);
                    else
                        memset(dst, 0, sizeof(or_cf64) * 2 * (size_t)nchan);
                }
            }
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int j = 0; j < nant; j++) {
                int b = j * nant - j * (j + 1) / 2;             /* first baseline (j, j+1), j-major */
                for (int kk = j + 1; kk < nant; kk++, b++) {
                    if (!valid[j] || !valid[kk]) continue;
                    for (int l = 0; l < 2; l++)
                        for (int m = 0; m < 2; m++) {
                            const or_cf64 *x = spec + ((size_t)j * 2 + l) * nchan;
                            const or_cf64 *yy = spec + ((size_t)kk * 2 + m) * nchan;
                            double *acc = vis + (((size_t)b * 4 + 2 * l + m) * nchan) * 2;
                            for (int ch = 0; ch < nchan; ch++) {
                                acc[2 * ch] += x[ch].re * yy[ch].re + x[ch].im * yy[ch].im;
                                acc[2 * ch + 1] += x[ch].im * yy[ch].re - x[ch].re * yy[ch].im;
                            }
                        }
                    count[b]++;
                }
                if (autos && valid[j]) {
                    for (int l = 0; l < 2; l++)
                        for (int m = 0; m < 2; m++) {
                            const or_cf64 *x = spec + ((size_t)j * 2 + l) * nchan;
                            const or_cf64 *yy = spec + ((size_t)j * 2 + m) * nchan;
                            double *acc = autos + (((size_t)j * 4 + 2 * l + m) * nchan) * 2;
                            for (int ch = 0; ch < nchan; ch++) {
                                acc[2 * ch] += x[ch].re * yy[ch].re + x[ch].im * yy[ch].im;
                                acc[2 * ch + 1] += x[ch].im * yy[ch].re - x[ch].re * yy[ch].im;
                            }
                        }
                    if (autocount) autocount[j]++;
                }
            }
        }
    }
    free(l0); free(l1); free(rot); free(u); free(y); free(fr); free(spec); free(valid);
    return 0;
}

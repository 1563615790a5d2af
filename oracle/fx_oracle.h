/*
 * fx_oracle.h -- CPU oracle for the FX correlator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it, and only as the checker.
 *
 * Two restatements of the reference's IPP CPU path
 * (kernel.c:709-878, processAntennasAndBaselinePara + processBaselinePara):
 *
 *   - FP32 restatement: the reference's operation order and precision, op for
 *     op (FP64 delay/phase arithmetic, f64->f32 narrowing of the rotator
 *     arguments and of the fractional delay, FP32 rotator products, FP32 FFT,
 *     FP32 accumulate).  Intel IPP is absent, so two primitives are our own:
 *       * sin/cos of an f32 argument: evaluated in FP64 and rounded to f32
 *         (a correctly rounded f32 sincos in all but double-rounding ties).
 *         The reference calls ippsSinCos_32f_A11 (11 correct bits), whose bit
 *         patterns no available source pins.
 *       * the forward C2C FFT (ippsFFTFwd_CToC_32fc, no normalisation):
 *         iterative radix-2 DIT with twiddles computed in FP64 and rounded.
 *   - FP64 oracle: the same discrete decisions (integer offsets, validity) and
 *     the same rotator factorisation, with every floating operation in FP64
 *     and no narrowing.  This is the truth against which the reduced-precision
 *     (FP16 / FP8) X-engines are measured.
 *
 * Generalised channel counts: the reference accepts only square nchan
 * (kernel.c:31-35).  For any power-of-two nchan the oracle uses
 * stride = 2^floor(log2(nchan)/2), sub = 2*stride, nstep = F/sub fringe steps
 * and nfstep = nchan/stride fractional-delay steps.  For square nchan these
 * are exactly the reference's stride, substride and loop bounds.
 */
#ifndef FX_ORACLE_H
#define FX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float re, im; } or_cf32;
typedef struct { double re, im; } or_cf64;

/* Derived per-run sizes, kernel.c:13-39 (generalised as described above). */
typedef struct {
    int nchan;      /* channels kept (lower sideband half of the FFT)       */
    int F;          /* FFT length = 2*nchan (real sampling)                */
    int stride;     /* fine frac-delay block / coarse fringe step unit      */
    int sub;        /* fringe-rotator fine block = 2*stride                 */
    int nstep;      /* fringe coarse steps = F/sub                          */
    int nfstep;     /* frac-delay coarse steps = nchan/stride               */
    double Ts;      /* sample time = 1/(2*bw)                               */
    double lo;      /* LO frequency (Hz)                                    */
    double bw;      /* bandwidth (Hz)                                       */
    int fracLO;     /* (lo - (int)lo) > 1e-11                               */
    int iscomplex;  /* 0: real 2-bit samples (the reference's CPU path);
                       1: complex 8-bit samples, F = nchan, Ts = 1/bw        */
    int origin;     /* global FFT index of the first FFT of the data buffer:
                       o = (g - origin)*F - n.  0 (set by or_plan_init) is the
                       reference, whose buffer starts at FFT 0; a streaming
                       caller's subint buffers start at later FFTs.         */
} or_plan;

/* Returns 0 on success, -1 if nchan is not a power of two >= 4. */
int or_plan_init(or_plan *p, int nchan, double lo, double bw);

/* Complex 8-bit input (the COMPLEX keyword, allocDataHost.c:12-19; the
 * reference's CPU path rejects it, kernel.c:19-28, its GPU path unpacks it,
 * gkernel.cu:277-296).  The same semantics as the real path with F = nchan
 * complex samples per FFT, Ts = 1/bw, every FFT bin a channel (the band is
 * taken as [0, bw)), stride/sub/nstep/nfstep from nchan by the same rules. */
int or_plan_init_complex(or_plan *p, int nchan, double lo, double bw);

/* Bytes per antenna of a complex 8-bit dual-polarisation subint: numffts + 1
 * FFTs of nchan samples, 4 bytes each (pol 0 re, im, pol 1 re, im as int8;
 * gkernel.cu:279-287).  allocDataHost.c:21 gives half of this for nbit 8
 * complex (it counts one byte per complex sample and polarisation). */
long long or_subint_bytes_complex(int nchan, int numffts);

/* The F complex samples of both polarisations starting at sample `offset`
 * (byte 4*offset): x0 = pol 0, x1 = pol 1. */
void or_unpack8complex(const uint8_t *data, int offset, int F, or_cf32 *x0, or_cf32 *x1);

/* kernel.c:373-386 getStationDelay, FP64, same evaluation order. */
void or_station_delay(const double c[3], int fftindex, double *mean, double *a, double *b);

/* kernel.c:736-756: integer/fractional split, edge clamps, validity.
 * Returns 1 if the antenna is valid for this FFT. */
int or_offset_origin(double mean, double fileoffset, double Ts, int g, int origin, int F, int maxoffset,
                     int *offset, double *fracdelay);
int or_offset(double mean, double fileoffset, double Ts, int g, int F, int maxoffset,
              int *offset, double *fracdelay);

/* kernel.c:189-202 / 397-426: 2-bit real unpack of F samples starting at
 * sample `offset` of the packed stream; both polarisations. */
void or_unpack2bit(const uint8_t *data, int offset, int F, float *pol0, float *pol1);

/* Level table L = {-3.3359f, -1, +1, +3.3359f} (kernel.c:191-192). */
float or_level(int code);

/* kernel.c:428-497 fringe rotator (FP32 result, FP64 phases), length F. */
void or_fringe_rotator_f32(const or_plan *p, double a, double b, or_cf32 *rot);
/* FP64 version (no narrowing, FP64 sincos and products). */
void or_fringe_rotator_f64(const or_plan *p, double a, double b, or_cf64 *rot);

/* kernel.c:512-555 fractional-sample rotator, length nchan. */
void or_frac_rotator_f32(const or_plan *p, double fracdelay, or_cf32 *fr);
void or_frac_rotator_f64(const or_plan *p, double fracdelay, or_cf64 *fr);

/* Forward C2C DFT, unnormalised: X[m] = sum_n x[n] exp(-2 pi i n m / F). */
void or_fft_f32(int F, const or_cf32 *in, or_cf32 *out);
void or_fft_f64(int F, const or_cf64 *in, or_cf64 *out);

/* sin/cos of an f32 argument, FP64-evaluated and rounded to f32. */
void or_sincos_f32(float x, float *s, float *c);

/* One call of the hot path: kernel.c:709-827 processAntennasAndBaselinePara.
 *
 * Inputs
 *   nant, numffts, split        maxoffset = (numffts*split - 1)*F (kernel.c:730)
 *   delays[nant*3]              c0 c1 c2 per antenna (FFT-index units -> s)
 *   fileoffsets[nant]           seconds
 *   iteration[numffts]          global FFT indices
 *   data[nant]                  packed 2-bit streams
 * Outputs (accumulated, +=, like the reference)
 *   vis[nbl*4*nchan*2]          cross baselines j<k, j-major, products 2l+m
 *   count[nbl]
 * Optional outputs (may be NULL)
 *   autos[nant*4*nchan*2]       same-antenna products (accumulated)
 *   autocount[nant]
 *   chan[numffts*nant*2*nchan*2] frac-corrected spectra, channels < nchan
 *   valid[numffts*nant]
 *   offsets[numffts*nant], fracdelays[numffts*nant]
 * Returns 0 on success. */
int or_process_f32(const or_plan *p, int nant, int numffts, int split,
                   const double *delays, const double *fileoffsets, const int *iteration,
                   const uint8_t *const *data,
                   float *vis, int *count, float *autos, int *autocount,
                   float *chan, int *valid, int *offsets, double *fracdelays);

/* FP64 oracle, same interface with double outputs. */
int or_process_f64(const or_plan *p, int nant, int numffts, int split,
                   const double *delays, const double *fileoffsets, const int *iteration,
                   const uint8_t *const *data,
                   double *vis, int *count, double *autos, int *autocount,
                   double *chan, int *valid);

/* F-stage only, FP32 restatement, with OpenMP over antennas when built with
 * -fopenmp: spectra chan[numffts][nant][2][nchan] and valid[numffts][nant].
 * Used by the CPU baseline timing. */
int or_fstage_f32(const or_plan *p, int nant, int numffts, int split,
                  const double *delays, const double *fileoffsets, const int *iteration,
                  const uint8_t *const *data, float *chan, int *valid, int nthreads);

/* X-stage only (FP32, reference accumulation order per baseline): adds
 * sum over FFTs of chan_j,l * conj(chan_k,m) into vis/count for j<k. */
int or_xstage_f32(int nant, int nchan, int numffts, const float *chan, const int *valid,
                  float *vis, int *count, int nthreads);

/* Number of bytes per antenna per subint (allocDataHost.c:21) for 2-bit
 * real data with 2 polarisations. */
long long or_subint_bytes(int nchan, int numffts);

#ifdef __cplusplus
}
#endif
#endif

/*
 * kernel.h -- drop-in replacement for the reference's kernel.h
 * (twiG56/difx_tests kernel.h:1-136), implemented by libdifxfx.so on an
 * AMD Instinct MI355X.
 *
 * Every prototype below has the signature of the reference declaration it
 * replaces (line numbers in brackets), so a host program written against the
 * reference (main.c) compiles and links unchanged, without Intel IPP.
 *
 * Behaviour:
 *   - processAntennasAndBaselinePara [106-110] runs the whole subint on the
 *     GPU (delay model, unpack, fringe rotation, FFT, fractional delay,
 *     X-engine) and adds the visibilities and valid counts of the call into
 *     the caller's host arrays (+=), exactly as kernel.c:709-827 accumulates.
 *     channelised[j][p][0..nchan) / conjchannels[j][p] receive the last valid
 *     spectrum of each antenna, as the reference leaves them.
 *   - init_FxKernel [63-66] derives the same sizes as kernel.c:13-39.  nchan
 *     must be a power of two; non-square nchan (which kernel.c:31-35 rejects)
 *     is accepted with stride = 2^floor(log2(nchan)/2).
 *   - The fine-grained per-(antenna, FFT) functions operate on host arrays
 *     and run as single-shot GPU calls.
 *   - Fatal errors print to stderr and exit(1), as the reference does.
 *   - Not re-entrant (the reference is not either).
 */
#ifndef TESTS_DIFX_KERNEL_H
#define TESTS_DIFX_KERNEL_H

#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>
#include <math.h>

#include "vectordefs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* [20-57] per-call scratch of the CPU kernel; the FFT spec is opaque. */
typedef struct {
    f64 *subtoff;
    f64 *subtval;
    f64 *subxoff;
    f64 *subxval;
    f64 *subphase;
    f32 *subarg;
    f32 *subsin;
    f32 *subcos;

    f64 *steptoff;
    f64 *steptval;
    f64 *stepxoff;
    f64 *stepxval;
    f64 *stepphase;
    f32 *steparg;
    f32 *stepsin;
    f32 *stepcos;
    cf32 *stepcplx;
    cf32 *complexrotator;

    f32 *subfracsamparg;
    f32 *subfracsampsin;
    f32 *subfracsampcos;
    f32 *subchannelfreqs;

    f32 *stepfracsamparg;
    f32 *stepfracsampsin;
    f32 *stepfracsampcos;
    cf32 *stepfracsampcplx;
    f32 *stepchannelfreqs;
    cf32 *fracsamprotator;

    u8 *fftbuffer;
    vecFFTSpecC_cf32 *pFFTSpecC;
} FxKernel;

/* [59] 2-bit look-up table (one copy per translation unit, as in the reference) */
static cf32 lut2bit[256][4] __attribute__((unused));

void initLUT2bitReal(void);                                                     /* [61] */

void init_FxKernel(int *nant, int *nchan, int *nbit, double *lo, double *bw, struct timespec *starttime,
                   char *starttimestring, int *fftchannels, double *sampletime, int *stridesize,
                   int *substridesize, int *fractionalLoFreq, cf32 ***unpacked, cf32 ***channelised,
                   cf32 ***conjchannels, cf32 ***visibilities, int *nbaselines, int *baselineCount,
                   int split);                                                  /* [63-66] */

void unpack(u8 *inputdata, cf32 **unpacked, int offset, int nbit, int fftchannels);           /* [68] */

void getStationDelay(int antenna, int fftindex, double *meandelay, double *a, double *b,
                     double **delays);                                          /* [70] */

void unpackReal2bit(u8 *inputdata, cf32 **unpacked, int offset, int nsamp);    /* [72] */

void fringerotate(FxKernel *kernel, cf32 **unpacked, f64 a, f64 b, int substridesize, int stridesize,
                  double lo, int fftchannels, double fractionalLoFreq);        /* [74] */

void dofft(FxKernel *kernel, cf32 **unpacked, cf32 **channelised);             /* [76] */

void fracSampleCorrect(FxKernel *kernel, cf32 **channelised, f64 fracdelay, int stridesize,
                       int nchan);                                              /* [78] */

void conjChannels(cf32 **channelised, cf32 **conjchannels, int nchan);         /* [80] */

void saveVisibilities(int *nbaselines, int *nchan, cf32 ***visibilities, double *bw);       /* [82] */

void startTiming(struct timespec *starttime, char *starttimestring);           /* [84] */

void saveLog(long long *diff_ms, char *starttimestring);                       /* [86] */

void endTiming(struct timespec *starttime, long long *diff_ms);                /* [88] */

void processAntennas(FxKernel *kernel, int *nant, int *numffts, int *fftchannels, double *antFileOffsets,
                     double *sampletime, int *antValid, u8 **inputData, cf32 ***unpacked,
                     cf32 ***channelised, cf32 ***conjchannels, double **delays, int *nbit,
                     int *substridesize, int *stridesize, int *fractionalLoFreq, double *lo,
                     int *nchan);                                               /* [90-93] */

void processNormalize(int *i, int *baselineCount, cf32 ***visibilities, int *nchan,
                      cf32 ***visibilities_out);                                /* [95] */

void processBaseline(int *nant, int *antValid, cf32 ***channelised, cf32 ***conjchannels,
                     cf32 ***visibilities, int *nchan, int *baselineCount);    /* [97-98] */

void processAntennasAndBaseline(int *nant, int *numffts, int *fftchannels, double *antFileOffsets,
                                double *sampletime, u8 **inputData, cf32 ***unpacked, cf32 ***channelised,
                                cf32 ***conjchannels, double **delays, int *nbit, int *substridesize,
                                int *stridesize, int *fractionalLoFreq, double *lo, int *nchan,
                                cf32 ***visibilities, int *baselineCount, double *bw,
                                cf32 ***visibilities_out, int *baselineCount_out);       /* [100-104] */

void processAntennasAndBaselinePara(int *nant, int *numffts, int *fftchannels, double *antFileOffsets,
                                    double *sampletime, u8 **inputData, cf32 ***unpacked,
                                    cf32 ***channelised, cf32 ***conjchannels, double **delays, int *nbit,
                                    int *substridesize, int *stridesize, int *fractionalLoFreq, double *lo,
                                    int *nchan, cf32 ***visibilities, int *baselineCount, int *iteration,
                                    double *bw, cf32 ***visibilities_out, int *baselineCount_out,
                                    int split);                                 /* [106-110] */

void resetBeforeProcess(int *nbaselines, cf32 ***visibilities, int *nchan, int *baselineCount,
                        int split);                                             /* [112] */

void manageAntennasAndBaseline(int *nant, int *numffts, int *fftchannels, double *antFileOffsets,
                               double *sampletime, u8 **inputData, cf32 ***unpacked, cf32 ***channelised,
                               cf32 ***conjchannels, double **delays, int *nbit, int *substridesize,
                               int *stridesize, int *fractionalLoFreq, double *lo, int *nchan,
                               cf32 ***visibilities, int *baselineCount, double *bw,
                               cf32 ***visibilities_out, int *baselineCount_out);        /* [114-118] */

void processBaselinePara(int *nant, int *antValid, cf32 ***channelised, cf32 ***conjchannels,
                         cf32 ***visibilities, int *nchan, int *baselineCount);         /* [120-121] */

void merge(int split, int nbaselines, cf32 ***visibilities, int *nant, int *nChan, int *baselineCount,
           cf32 ***visibilities_out, int *baselineCount_out);                  /* [123] */

void unpackImpl(cf32 **unpacked, int split, int *nant, int *fftchannels, u8 *inputData, int *offset,
                int *nbit);                                                     /* [125] */
void fringeRotateImpl(FxKernel *kernel, cf32 **unpacked, f64 a, f64 b, int substridesize, int stridesize,
                      double lo, int fftchannels, double fractionalLoFreq, cf32 **unpacked_out,
                      FxKernel *kernel_out);                                    /* [126-127] */
void doFFTImpl(FxKernel *kernel, cf32 **unpacked, cf32 **channelised, int *fftchannels);     /* [128] */
void fracSampleCorrectImpl(FxKernel *kernel, cf32 **channelised, double *fractionaldelay, int *stridesize,
                           int *nchan, cf32 **channelised_out);                 /* [129-130] */
void conjChannelsImpl(cf32 **channelised, cf32 **conjchannels, int *nchan, int *fftchannels,
                      cf32 **channelised_out);                                  /* [131] */
void stationDelayAndOffset(double **delays, int *iteration, double *antFileOffsets, int *antValid,
                           int *offset, double *fracDelay, double *delaya, double *delayb,
                           double *sampletime, int *fftchannels, int numffts);  /* [132-134] */

#ifdef __cplusplus
}
#endif
#endif /* TESTS_DIFX_KERNEL_H */

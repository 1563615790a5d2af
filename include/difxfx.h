/*
 * difxfx.h -- native C API of the MI355X FX correlator engine.
 *
 * Plain C ABI: pointers, sizes and status codes, no library types in the
 * signatures.  Every function returns 0 on success and a negative code on
 * failure (fx_last_error() gives the message).  A context owns the device
 * buffers of one subband and one HIP stream; calls on one context must not
 * overlap (the reference path is single threaded as well).
 *
 * The kernel.h entry points of the reference (kernel.h:61-134) are layered
 * on top of this API; see kernel.h in this directory.
 *
 * Pipeline of one call (reference: processAntennasAndBaselinePara,
 * kernel.c:709-827):
 *   fx_fstage   FP64 delay table -> fused unpack / fringe rotation / FFT /
 *               fractional-delay kernel -> corner-turned spectra in HBM
 *   fx_xstage   MFMA X-engine: accumulate the Hermitian visibility matrix
 *               of every channel over a range of FFT slots
 *   fx_read_vis reference-layout cross visibilities [b][2l+m][chan], the
 *               valid-FFT counts per baseline, and the autocorrelations
 */
#ifndef DIFXFX_H
#define DIFXFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_PREC_F32 0   /* spectra and X-engine in FP32 (plain path)            */
#define FX_PREC_F16 1   /* spectra cast to FP16, MFMA with FP32 accumulation     */
#define FX_PREC_F8  2   /* spectra scaled per (channel, signal) to FP8 e4m3      */
#define FX_PREC_F32V 3  /* FP32 spectra, plain FP32 vector-ALU X-engine (no MFMA) */

typedef struct fx_context fx_context;

typedef struct {
    int nant;                    /* antennas (2 polarisations each)                */
    int nchan;                   /* channels; FFT length F = 2*nchan, power of two */
    double lo;                   /* LO frequency, Hz (< 2^31, kernel.c:39)         */
    double bw;                   /* bandwidth, Hz; sample time 1/(2*bw)            */
    int prec;                    /* FX_PREC_*                                      */
    int max_ffts;                /* FFT slots held for the X-engine (device memory
                                    per slot and signal: the X input, nchan x 8, 4
                                    or 2 bytes, and at nchan = 16384 the F-stage's
                                    row staging, 64, 48 or 28 KB for FP32, FP16,
                                    FP8)                                          */
    long long max_bytes_per_ant; /* packed input capacity per antenna              */
    int iscomplex;               /* 0: real 2-bit samples, F = 2*nchan (the
                                    reference's path); 1: complex 8-bit samples,
                                    F = nchan, Ts = 1/bw, every FFT bin a channel
                                    (band taken as [0, bw)); a time sample is 4
                                    bytes: pol 0 re, im, pol 1 re, im as int8
                                    (gkernel.cu:279-287)                          */
} fx_params;

/* Status codes */
#define FX_OK 0
#define FX_EINVAL -1
#define FX_EHIP -2
#define FX_ENOMEM -3
#define FX_EUNSUPPORTED -4

int fx_create(const fx_params *prm, fx_context **out);
void fx_destroy(fx_context *ctx);
const char *fx_last_error(void);

/* FFT slots are grouped: fx_fstage writes whole X-engine K chunks of this
 * many FFTs. */
int fx_fft_granule(const fx_context *ctx);

/* Delay model: delays[3*a + {0,1,2}] = c0, c1, c2 (seconds, t in FFT units;
 * kernel.c:378), fileoffsets[a] in seconds (loadConfig.h:30-31). */
int fx_set_model(fx_context *ctx, const double *delays, const double *fileoffsets);

/* Packed 2-bit input, host -> device.  data[a] holds nbytes bytes. */
int fx_upload_data(fx_context *ctx, const uint8_t *const *data, long long nbytes);
/* Same from one host buffer with antenna a at data + a*stride. */
int fx_upload_data_strided(fx_context *ctx, const uint8_t *data, long long stride, long long nbytes);
/* Device address and per-antenna stride of the packed input buffer. */
int fx_device_data(fx_context *ctx, void **dptr, long long *stride);

/* F-stage for numffts FFTs with global indices iteration[0..numffts)
 * (host array), maximum offset (numffts*split - 1)*F (kernel.c:730), writing
 * FFT slots [fft0, fft0 + padded numffts); fft0 must be a multiple of the
 * granule.  *nslots receives the number of slots written. */
int fx_fstage(fx_context *ctx, int numffts, const int *iteration, int split, int fft0, int *nslots);
/* fx_fstage on input slot `slot` (0: the buffer of fx_upload_data; 1: the
 * second streaming slot) for a buffer whose first sample is that of global
 * FFT `origin`: the sample offset of FFT g is (g - origin)*F - n, with the
 * edge clamps and validity of kernel.c:742-756 applied at the buffer's own
 * edges.  origin 0 is the reference (one subint buffer starting at FFT 0);
 * a streaming caller passes global FFT indices and origin = the subint's
 * first FFT, so delay polynomials keep one time origin across subints. */
int fx_fstage_ex(fx_context *ctx, int slot, int numffts, const int *iteration, int split, int origin, int fft0,
                 int *nslots);

/* Streaming ingest: two input slots, each a device buffer of the context's
 * input capacity and a pinned host staging buffer of the same layout.
 *   fx_ingest_host    pinned host memory of a slot; antenna a's bytes go at
 *                     ptr + a*stride (the caller fills it, e.g. from a file)
 *   fx_ingest_submit  asynchronous H2D of the slot on the context's copy
 *                     stream, after the last F-stage that read the slot
 *   fx_ingest_wait    blocks until the slot's copy is complete, so its host
 *                     buffer may be refilled
 * An F-stage on a slot (fx_fstage_ex) waits for that slot's latest copy.
 * With two slots the copy of subint s+1 overlaps the F-stage and X-engine of
 * subint s. */
int fx_ingest_host(fx_context *ctx, int slot, uint8_t **ptr, long long *stride);
int fx_ingest_submit(fx_context *ctx, int slot, long long nbytes);
int fx_ingest_wait(fx_context *ctx, int slot);

/* X-engine over FFT slots [fft0, fft0 + nslots), accumulated into the
 * context's visibilities. */
int fx_xstage(fx_context *ctx, int fft0, int nslots);
/* fx_fstage at slot 0 followed by fx_xstage over the slots it wrote, as one
 * graph launch (fx_subint). */
int fx_process(fx_context *ctx, int numffts, const int *iteration, int split);

/* One subintegration as ONE hipGraph launch: the delay / rotator kernel and
 * the F-stage of fx_fstage_ex(slot, numffts, iteration, split, origin, fft0)
 * and, when xslots > 0, the X-engine and valid-FFT counts of
 * fx_xstage(0, xslots) after them (the X-engine on batch boundaries: a caller
 * batching k subints passes xslots = 0 for the first k-1 and the total slot
 * count with the last).  The graph is recorded by stream capture on the first
 * call of a shape (numffts, split, slot, fft0, xslots and the tuning options)
 * and replayed after that; FFT indices and the origin are copied to the
 * device ahead of each launch, so one recording serves every subint of a
 * stream.  Results equal the plain launches bit for bit; with FX_OPT_GRAPH 0
 * the same work goes out as plain launches.  The kernel.h entry point and
 * fx_process use this.  *nslots (may be NULL) receives the slots written. */
int fx_subint(fx_context *ctx, int slot, int numffts, const int *iteration, int split, int origin, int fft0,
              int xslots, int *nslots);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  fx_process followed by fx_read_vis of one subint,
 * without a context: every input and output is the caller's device memory
 * (packed samples, antenna a at data + a*stride; delays [nant][3],
 * fileoffsets [nant] and iteration [numffts] on the device), and every
 * intermediate lives in one file-scope device array of FX_NOALLOC_SCRATCH_MB
 * MiB per GPU, held until the process ends; nothing is allocated or freed.
 * Outputs (any may be NULL) are overwritten with this call's unnormalised
 * visibilities in fx_read_vis's layout: cross [nbl][4][nchan][2],
 * cross_counts [nbl], autos [nant][4][nchan][2], auto_counts [nant].  The
 * work is queued on `stream` (a hipStream_t; NULL is the default stream) and
 * the call returns without waiting.  Default radix plan and X-engine tile;
 * FP8 takes the analytic default scales.  data must be 16-byte aligned and
 * stride a multiple of 16 of at least the subint's bytes + 16.  A workload
 * whose accumulators and one granule of FFTs do not fit the array fails with
 * FX_ENOMEM before any launch (the accumulators take 8 KiB per channel and
 * pair of 32-signal blocks: 64 MiB holds up to 16 antennas at 4096 channels,
 * 48 at 1024); larger FFT counts go through in chunks.  Every call on a GPU shares that GPU's array, which keeps the last
 * call's contents: calls that overlap on one GPU (two streams, two threads)
 * race. */
#define FX_NOALLOC_SCRATCH_MB 64
int fx_process_noalloc(const fx_params *prm, const uint8_t *data, long long stride, const double *delays,
                       const double *fileoffsets, const int *iteration, int numffts, int split, float *cross,
                       int *cross_counts, float *autos, int *auto_counts, void *stream);

/* Zero the accumulated visibilities and counts. */
int fx_reset(fx_context *ctx);

/* Copy out the accumulated (unnormalised) visibilities:
 *   cross[((b*4 + p)*nchan + c)*2 + {re,im}]  b over j<k j-major, p = 2l+m
 *   cross_counts[b]  FFTs where both antennas were valid
 *   autos[((a*4 + p)*nchan + c)*2 + {re,im}], auto_counts[a]
 * Any output pointer may be NULL. */
int fx_read_vis(fx_context *ctx, float *cross, int *cross_counts, float *autos, int *auto_counts);
/* Same with flags: FX_VIS_NORMALIZE divides every visibility by its
 * valid-FFT count on the device (processNormalize, kernel.c:348-359;
 * zero counts leave the visibility as it is). */
#define FX_VIS_NORMALIZE 1
int fx_read_vis_ex(fx_context *ctx, int flags, float *cross, int *cross_counts, float *autos, int *auto_counts);

/* Visibility files of the accumulated visibilities (flags as fx_read_vis_ex).
 * FX_VIS_TEXT    the reference's vis.out (kernel.c:570-630) of the cross
 *                visibilities, one line per channel
 * FX_VIS_BINARY  "DIFXFXV1", int32 nant, nchan, nbl, flags, double bw, lo,
 *                int32 cross_counts[nbl], int32 auto_counts[nant],
 *                float cross[nbl][4][nchan][2], float autos[nant][4][nchan][2]
 *                (little endian; for large antenna counts, where the text
 *                file would run to gigabytes) */
#define FX_VIS_TEXT 0
#define FX_VIS_BINARY 1
int fx_save_vis(fx_context *ctx, const char *path, int format, int flags);

/* Debug taps (host copies).  Delay table entries are 32 bytes:
 * {double a, b, fd; int offset, valid} per (FFT, antenna) of the last
 * fx_fstage call. */
int fx_read_delay_table(fx_context *ctx, int nffts, void *out);
/* When enabled, fx_fstage also writes FP32 spectra [fft][ant][pol][nchan]
 * (complex) of its FFTs; fx_read_tap copies nffts of them.  enable is a mask
 * of FX_TAP_SPECTRA and FX_TAP_LEVELS (levels imply spectra); FX_TAP_LEVELS
 * also keeps the 2-bit levels the fused F-stage decoded for every sample of
 * the valid FFTs (real input), [fft][ant][pol][F] floats, copied by
 * fx_read_tap_levels (kernel.c:397-426 unpack, inside the fused kernel).
 * Entries of invalid (FFT, antenna) pairs are left unwritten. */
#define FX_TAP_SPECTRA 1
#define FX_TAP_LEVELS 2
int fx_enable_tap(fx_context *ctx, int enable);
int fx_read_tap(fx_context *ctx, int nffts, float *out);
int fx_read_tap_levels(fx_context *ctx, int nffts, float *out);

/* FP8: per (channel, signal) scales, scales[c*2*nant + 2*a + pol]; the
 * stored value is spectrum/scale.  Defaults to the analytic 2-bit power of a
 * flat spectrum (one rms per component -> 16).  fx_update_fp8_scales sets
 * them on the device from the autocorrelations accumulated since fx_reset:
 * per (channel, signal) the rms of each component, sqrt(P / 2n) for auto
 * power P over n valid FFTs, maps to 16 (signals without a valid FFT keep
 * their scale).  With FX_OPT_FP8_AUTOSCALE 1 this follows every X-engine
 * launch, so each batch of F-stages is scaled by the autocorrelations of the
 * subints before it (a sloped bandpass or a strong narrow-band tone no longer
 * saturates e4m3 at 448 after the first batch of an integration; scales carry
 * over fx_reset into the next integration).  The X-engine applies s_i s_j
 * before accumulating, so scales may change between its launches; a caller
 * must not run F-stages of the next batch before the X-engine of the
 * current one. */
int fx_set_fp8_scales(fx_context *ctx, const float *scales);
int fx_update_fp8_scales(fx_context *ctx);

int fx_sync(fx_context *ctx);

/* Kernel timing with HIP events on the context's stream: when enabled, every
 * fx_fstage / fx_xstage records events and fx_get_timing returns the summed
 * milliseconds since the last fx_reset_timing of {delay, fstage, counts,
 * xstage} and the launch counts of {fstage, xstage}. */
int fx_set_timing(fx_context *ctx, int enable);
int fx_reset_timing(fx_context *ctx);
int fx_get_timing(fx_context *ctx, double *ms4, int *nlaunch2);
/* Graph launches of fx_subint are timed as a whole (events around each
 * launch): summed milliseconds and launch count since fx_reset_timing. */
int fx_get_graph_timing(fx_context *ctx, double *ms, int *nlaunch);

/* Geometry used by the kernels (for reports): fills
 * {F, stride, sub, nstep, nfstep, S, Spad, nblk, nbp, G, granule}. */
int fx_get_geometry(const fx_context *ctx, int *out11);

/* Options.  The tuning options (the autotuner sets them) leave results
 * unchanged; FX_OPT_FP8_AUTOSCALE does not.
 * FX_OPT_XTILE   X-engine workgroup tile, a square of 1, 2 or 4 blocks of
 *                32x32 signals per panel; 0 selects by the signal count
 * FX_OPT_FPLAN   F-stage radix plan, 0 .. FX_OPT_NPLANS - 1
 * FX_OPT_NPLANS  (read only) number of radix plans at this FFT length and input
 * FX_OPT_XBATCH  subints per X-engine launch recommended to callers that
 *                batch fx_fstage calls before one fx_xstage (default 1)
 * FX_OPT_GRAPH   1 (default): fx_subint / fx_process launch recorded graphs;
 *                0: the same kernels as plain launches
 * FX_OPT_FP8_AUTOSCALE  1: FP8 scales follow the accumulated
 *                autocorrelations after every X-engine launch (see
 *                fx_update_fp8_scales); 0 (default): fixed scales.  FP8 only;
 *                unlike the others this option changes results */
#define FX_OPT_XTILE 1
#define FX_OPT_FPLAN 2
#define FX_OPT_NPLANS 3
#define FX_OPT_XBATCH 4
#define FX_OPT_GRAPH 5
#define FX_OPT_FP8_AUTOSCALE 6
int fx_set_option(fx_context *ctx, int option, int value);
int fx_get_option(const fx_context *ctx, int option, int *value);

/* Autotuner: times every F-stage radix plan, every X-engine tile and 1, 2,
 * 4, ... subints per X-engine launch (up to the FFT capacity) on this
 * context's uploaded input and delay model, numffts FFTs per subint, each
 * candidate once untimed then `reps` times (in three round-robin rounds,
 * best round kept); keeps the fastest plan, then times the tiles within 25%
 * of the fastest tile alone together with each launch count as whole steps of
 * that many F-stages and one X-engine launch, and keeps the smallest count
 * whose best tile is within 1% of the best time per subint, with that tile;
 * zeroes the visibilities and counts.  report (may be NULL) receives {plan,
 * tile, subints per launch, F-stage ms per subint, X-engine ms per subint
 * alone (up to 8 subints per launch)}.  When
 * cache_path (or, if it is NULL, the FX_TUNE_CACHE environment variable) names
 * a file, the result is stored there per (GPU arch, nant, nchan, prec), and
 * fx_create applies a stored entry when FX_TUNE_CACHE names that file. */
int fx_autotune(fx_context *ctx, int numffts, int reps, const char *cache_path, double *report);

/* Batches: several contexts (one per subband, each with its own antennas,
 * channels, bandwidth and precision) processed one subint at a time.
 * fx_batch_process runs fx_process(ctx[i], numffts[i], iteration[i], split)
 * for every context:
 *   FX_BATCH_GRAPH   as one hipGraph launch per call, the subbands as
 *                    concurrent branches; the graph is recorded on the first
 *                    call and again only when numffts, split or a context's
 *                    options change
 *   FX_BATCH_LAUNCH  as plain launches, one subband after another
 * Work queued earlier on a context (uploads, fx_reset) runs first; later calls
 * on a context run after the subint.  Results are identical in both modes. */
typedef struct fx_batch fx_batch;
#define FX_BATCH_LAUNCH 0
#define FX_BATCH_GRAPH 1
int fx_batch_create(fx_context *const *ctxs, int n, fx_batch **out);
void fx_batch_destroy(fx_batch *batch);
int fx_batch_process(fx_batch *batch, const int *numffts, const int *const *iteration, int split, int mode);
/* Same with a buffer origin per context (fx_fstage_ex); origins may be NULL
 * (all 0).  Origins are read by the kernels from device memory, so one
 * recorded graph serves every subint of a stream. */
int fx_batch_process_ex(fx_batch *batch, const int *numffts, const int *const *iteration, const int *origins,
                        int split, int mode);
int fx_batch_sync(fx_batch *batch);
/* Number of graph recordings so far (for tests). */
int fx_batch_captures(const fx_batch *batch);

#ifdef __cplusplus
}
#endif
#endif

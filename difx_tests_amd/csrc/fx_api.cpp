// fx_api.cpp -- context management and the native C API (include/difxfx.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "difxfx.h"
#include "fx_common.h"
#include "fx_kernels.h"

using namespace fx;

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(FX_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                              \
    } while (0)

// Kernel timing: event pairs recorded around launches on the context's
// stream without synchronising; fx_get_timing waits and sums them.
enum { TM_DELAY, TM_FSTAGE, TM_COUNTS, TM_XSTAGE, TM_GRAPH, TM_N };
struct EvPair {
    hipEvent_t a, b;
    int kind;
};

struct fx_context {
    FxPlan p;
    int prec;           // X-engine precision (FX_PREC_*)
    int fprec;          // F-stage output format (FP32 for both FP32 X-engines)
    int G;              // FFTs per corner-turn group
    int kgroups;        // groups per X-engine K chunk
    int granule;        // G * kgroups: FFT slots per K chunk
    int nfg;            // groups held by the X input
    int max_ffts;       // padded FFT capacity
    long long cap_bytes, stride;
    hipStream_t stream;
    uint8_t* d_data;    // input slot 0 (also the buffer of fx_upload_data / fx_fstage)
    uint8_t* d_data1;   // input slot 1 (streaming ingest), allocated on first use
    uint8_t* h_pin[2];  // pinned host staging of the two slots, allocated on first use
    hipStream_t cstream;              // copy stream of the streaming ingest
    hipEvent_t copied[2], slot_free[2];
    int ingest;         // streaming ingest in use: F-stages wait for their slot's copy
    double* d_model;    // delays [nant*3] then fileoffsets [nant]
    int* d_iter;
    FxDelay* d_dtab;
    uint8_t* d_valid;   // [max_ffts][nant] validity of each FFT slot's antennas (counts)
    float2* d_rtab;     // rotator tables [max_ffts][nant][sub + nstep + stride + nfstep]
    float2* d_tw;
    float* d_subchf;
    float* d_stepchf;
    void* d_xin;
    size_t xin_bytes;
    float2* d_vacc;
    size_t vacc_bytes;
    int* d_counts;
    float* d_scale;     // [nchan][Spad] (FP8, X-engine)
    float* d_scale_inv; // [Spad][nchan] (FP8, F-stage)
    float2* d_cross;
    float2* d_autos;
    float2* d_tap;
    float* d_ltap;      // decoded levels of the tap (FX_TAP_LEVELS)
    void* d_stage;      // F = 32768: the F-stage's row staging (fx_fstage_stage_bytes)
    int tap;            // FX_TAP_* bits
    int last_nffts;     // entries in the delay table of the last fx_fstage
    // pinned ring of {origin, first FFT index, list flag} blocks copied to
    // d_iter + max_ffts ahead of launches that read them from the device
    int* h_ring;
    hipEvent_t ring_ev[64];
    unsigned char ring_used[64];
    int ring_pos;
    int xsb;            // X-engine super-block size (0: by the signal count)
    int fplan;          // F-stage radix plan
    int xbatch;         // subints per X-engine launch recommended to callers (1 unless tuned)
    int timing;
    int capturing;      // launches are being recorded into a graph: no timing events
    std::vector<EvPair> pool;      // free event pairs
    std::vector<EvPair> pending;   // recorded, not yet summed
    double ms[TM_N];
    int nl[3];          // launches: F-stage, X-engine, subint graph
    // recorded subint graphs (fx_subint), one per call shape, each
    // instantiated kExecCopies times
    struct SubGraph {
        long long key[10];
        hipGraph_t graph;          // kept: its delay-kernel node addresses argument updates
        hipGraphNode_t dnode;      // delay / rotator kernel node (arguments updated per launch), or null
        hipKernelNodeParams kp;    // its recorded parameters
        hipGraphExec_t exec[8];
        hipEvent_t done[8];        // each copy's last launch
        unsigned char used[8];
        int next;
    };
    std::vector<SubGraph> graphs;
    int graph_args = 1;  // recorded graphs take contiguous FFT indices as updated kernel arguments
    int use_graph = 1;  // FX_OPT_GRAPH
    int autoscale = 0;  // FX_OPT_FP8_AUTOSCALE
};

static constexpr int kRing = 64;
// FP8: one rms of each spectral component maps to this value (e4m3 holds up to
// 448, so a single FFT's component saturates only beyond 28 rms)
static constexpr float kFp8Target = 16.0f;

static constexpr int kExecCopies = 8;
static_assert(sizeof(((fx_context::SubGraph*)nullptr)->exec) == kExecCopies * sizeof(hipGraphExec_t),
              "SubGraph holds kExecCopies executable graphs");

static void subgraph_free(fx_context::SubGraph& g)
{
    for (int i = 0; i < kExecCopies; i++) {
        if (g.exec[i]) (void)hipGraphExecDestroy(g.exec[i]);
        if (g.done[i]) (void)hipEventDestroy(g.done[i]);
    }
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = fx_context::SubGraph{};
}

static int xin_row_bytes(int prec)
{
    (void)prec;
    return 16;  // F32: (re0, re1, im0, im1); F16: 4 x half2; F8: 8 x 2 fp8
}

extern "C" const char* fx_last_error(void) { return g_err; }

extern "C" int fx_fft_granule(const fx_context* ctx) { return ctx ? ctx->granule : FX_EINVAL; }

static void free_ctx(fx_context* c)
{
    if (c->cstream) (void)hipStreamSynchronize(c->cstream);
    for (int k = 0; k < 2; k++) {
        if (c->h_pin[k]) (void)hipHostFree(c->h_pin[k]);
        if (c->copied[k]) (void)hipEventDestroy(c->copied[k]);
        if (c->slot_free[k]) (void)hipEventDestroy(c->slot_free[k]);
    }
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    void* bufs[] = {c->d_data, c->d_data1, c->d_model, c->d_iter, c->d_dtab, c->d_valid, c->d_rtab, c->d_tw, c->d_subchf, c->d_stepchf,
                    c->d_xin, c->d_vacc, c->d_counts, c->d_scale, c->d_scale_inv,
                    c->d_cross, c->d_autos, c->d_tap, c->d_ltap, c->d_stage};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (auto* v : {&c->pool, &c->pending})
        for (EvPair& e : *v) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
    for (auto& g : c->graphs) subgraph_free(g);
    for (int i = 0; i < kRing; i++)
        if (c->ring_ev[i]) (void)hipEventDestroy(c->ring_ev[i]);
    if (c->h_ring) (void)hipHostFree(c->h_ring);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" void fx_destroy(fx_context* ctx)
{
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
}

static void tune_apply_cached(fx_context* c);

static int make_plan(const fx_params* prm, FxPlan* p);

namespace fx {
int plan_from_params(const fx_params* prm, FxPlan* p) { return make_plan(prm, p); }
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace fx

// Host-side plan: kernel.c:13-39, generalised to power-of-two nchan.
static int make_plan(const fx_params* prm, FxPlan* p)
{
    memset(p, 0, sizeof(*p));
    if (prm->iscomplex != 0 && prm->iscomplex != 1) return fail(FX_EINVAL, "iscomplex=%d", prm->iscomplex);
    const int cplx = prm->iscomplex;
    const int F = cplx ? prm->nchan : 2 * prm->nchan;
    const int lg = fx_ilog2_exact(F);
    if (lg < 0) return fail(FX_EINVAL, "nchan=%d is not a power of two", prm->nchan);
    if (cplx ? !fstage_supported_complex(lg) : !fstage_supported(lg))
        return fail(FX_EUNSUPPORTED, "%s FFT length %d not supported (%s)", cplx ? "complex" : "real", F,
                    cplx ? "512..16384" : "512..32768");
    if (prm->nant < 1) return fail(FX_EINVAL, "nant=%d", prm->nant);
    if (!(prm->bw > 0)) return fail(FX_EINVAL, "bandwidth must be positive");
    if (!(prm->lo >= 0 && prm->lo < 2147483647.0)) return fail(FX_EINVAL, "LO must be in [0, 2^31) Hz");
    p->nant = prm->nant;
    p->nchan = prm->nchan;
    p->F = F;
    p->lg = lg;
    p->iscomplex = cplx;
    p->stride = 1 << (fx_ilog2_exact(prm->nchan) / 2);   // 2^floor(log2(nchan)/2)
    p->sub = 2 * p->stride;
    p->nstep = p->F / p->sub;
    p->nfstep = p->nchan / p->stride;
    p->S = 2 * prm->nant;
    p->Spad = (p->S + FX_BLK - 1) / FX_BLK * FX_BLK;
    p->nblk = p->Spad / FX_BLK;
    p->nbp = p->nblk * (p->nblk + 1) / 2;
    p->Ts = cplx ? 1.0 / prm->bw : 1.0 / (2.0 * prm->bw);
    p->lo = prm->lo;
    p->bw = prm->bw;
    p->lofrac = prm->lo - (double)(int)prm->lo;
    p->fracLO = p->lofrac > 0.00000000001;
    return FX_OK;
}

extern "C" int fx_create(const fx_params* prm, fx_context** out)
{
    if (!prm || !out) return fail(FX_EINVAL, "null argument");
    *out = nullptr;
    if (prm->prec < FX_PREC_F32 || prm->prec > FX_PREC_F32V) return fail(FX_EINVAL, "bad precision %d", prm->prec);
    if (prm->max_ffts < 1) return fail(FX_EINVAL, "max_ffts must be >= 1");
    FxPlan plan;
    int rc = make_plan(prm, &plan);   // parameters are checked before any device call
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(FX_EHIP, "no HIP device available");
    fx_context* c = new fx_context();
    c->p = plan;
    FxPlan& p = c->p;
    c->prec = prm->prec;
    c->fprec = prm->prec == FX_PREC_F32V ? FX_PREC_F32 : prm->prec;   // F-stage output format
    c->xbatch = 1;
    c->G = fx_fstage_group_size(c->fprec);
    c->kgroups = xengine_chunk_groups(prm->prec);
    c->granule = c->G * c->kgroups;
    p.G = c->G;
    c->max_ffts = (prm->max_ffts + c->granule - 1) / c->granule * c->granule;
    c->nfg = c->max_ffts / c->G;
    // allocDataHost.c:21 for real 2-bit data; complex 8-bit data take 4 bytes
    // per sample (both polarisations' int8 re, im)
    const long long need = p.iscomplex ? 4LL * p.F * (prm->max_ffts + 1) : (long long)p.nchan * (prm->max_ffts + 1);
    c->cap_bytes = prm->max_bytes_per_ant > need ? prm->max_bytes_per_ant : need;
    c->stride = (c->cap_bytes + 64 + 255) / 256 * 256;

#define ALLOC(ptr, bytes)                                                                   \
    do {                                                                                    \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                             \
            free_ctx(c);                                                                    \
            return fail(FX_ENOMEM, "hipMalloc of %zu bytes failed", (size_t)(bytes));       \
        }                                                                                   \
    } while (0)

    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        free_ctx(c);
        return fail(FX_EHIP, "hipStreamCreate failed");
    }
    ALLOC(c->d_data, (size_t)c->stride * p.nant);
    ALLOC(c->d_model, sizeof(double) * 4 * p.nant);
    ALLOC(c->d_iter, sizeof(int) * (c->max_ffts + 4));   // FFT indices, then {origin, it0, list flag}
    if (hipHostMalloc((void**)&c->h_ring, sizeof(int) * 4 * kRing, hipHostMallocDefault) != hipSuccess) {
        free_ctx(c);
        return fail(FX_ENOMEM, "pinned FFT-index ring failed");
    }
    for (int i = 0; i < kRing; i++)
        if (hipEventCreateWithFlags(&c->ring_ev[i], hipEventDisableTiming) != hipSuccess) {
            free_ctx(c);
            return fail(FX_EHIP, "hipEventCreate failed");
        }
    ALLOC(c->d_dtab, sizeof(FxDelay) * (size_t)c->max_ffts * p.nant);
    ALLOC(c->d_valid, (size_t)c->max_ffts * p.nant);
    ALLOC(c->d_rtab, sizeof(float2) * (size_t)c->max_ffts * p.nant * fx_rtab_len(p));
    if (const size_t sb = fx_fstage_stage_bytes(p.lg, c->fprec, p.nant, c->nfg)) ALLOC(c->d_stage, sb);
    ALLOC(c->d_tw, sizeof(float2) * p.F);
    ALLOC(c->d_subchf, sizeof(float) * p.stride);
    ALLOC(c->d_stepchf, sizeof(float) * p.nfstep);
    c->xin_bytes = (size_t)(c->nfg / c->kgroups) * p.nchan * fx_xin_pitch(c->kgroups, p.Spad) * xin_row_bytes(c->prec);
    ALLOC(c->d_xin, c->xin_bytes);
    c->vacc_bytes = (size_t)p.nchan * p.nbp * FX_BLK * FX_BLK * sizeof(float2);
    ALLOC(c->d_vacc, c->vacc_bytes);
    ALLOC(c->d_counts, sizeof(int) * p.nant * p.nant);
    ALLOC(c->d_scale, sizeof(float) * (size_t)p.nchan * p.Spad);
    ALLOC(c->d_scale_inv, sizeof(float) * (size_t)p.nchan * p.Spad);
    const int nbl = p.nant * (p.nant - 1) / 2;
    ALLOC(c->d_cross, sizeof(float2) * ((size_t)nbl > 0 ? (size_t)nbl : 1) * 4 * p.nchan);
    ALLOC(c->d_autos, sizeof(float2) * (size_t)p.nant * 4 * p.nchan);

    // twiddles and channel-frequency tables (FP64 evaluated, rounded to f32)
    std::vector<float2> tw(p.F);
    for (int m = 0; m < p.F; m++) {
        const double ang = -2.0 * M_PI * (double)m / (double)p.F;
        tw[m] = make_float2((float)cos(ang), (float)sin(ang));
    }
    std::vector<float> subchf(p.stride), stepchf(p.nfstep);
    const double PI = 3.14159265358979323846;
    for (int s = 0; s < p.stride; s++) subchf[s] = (float)((2 * PI * s * p.bw) / (double)p.nchan);     // kernel.c:164
    for (int k = 0; k < p.nfstep; k++)
        stepchf[k] = (float)((2 * PI * k * p.stride * p.bw) / (double)p.nchan);                      // kernel.c:165
#undef ALLOC
    // FP8 default scales: the spectrum of 2-bit noise has E|Y|^2 = F * E[L^2]
    // with E[L^2] = 4.3128 for levels {+-1, +-3.3359} at +-0.98 sigma; map one
    // standard deviation of each component to 16.
    // complex input: the default assumes int8 components at 24 rms
    const float sigma = (float)sqrt(p.iscomplex ? p.F * 2.0 * 24.0 * 24.0 / 2.0 : p.F * 4.3128 / 2.0);
    std::vector<float> sc((size_t)p.nchan * p.Spad, sigma / 16.0f), sci((size_t)p.nchan * p.Spad, 16.0f / sigma);

    if (hipMemcpy(c->d_tw, tw.data(), sizeof(float2) * p.F, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_subchf, subchf.data(), sizeof(float) * p.stride, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_stepchf, stepchf.data(), sizeof(float) * p.nfstep, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_scale, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_scale_inv, sci.data(), sizeof(float) * sci.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->d_data, 0, (size_t)c->stride * p.nant) != hipSuccess ||
        hipMemset(c->d_model, 0, sizeof(double) * 4 * p.nant) != hipSuccess ||
        hipMemset(c->d_xin, 0, c->xin_bytes) != hipSuccess ||
        hipMemset(c->d_vacc, 0, c->vacc_bytes) != hipSuccess ||
        hipMemset(c->d_counts, 0, sizeof(int) * p.nant * p.nant) != hipSuccess ||
        hipMemset(c->d_valid, 0, (size_t)c->max_ffts * p.nant) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        free_ctx(c);
        return fail(FX_EHIP, "context initialisation failed");
    }
    tune_apply_cached(c);
    *out = c;
    return FX_OK;
}

extern "C" int fx_set_model(fx_context* c, const double* delays, const double* fileoffsets)
{
    if (!c || !delays || !fileoffsets) return fail(FX_EINVAL, "null argument");
    std::vector<double> m(4 * (size_t)c->p.nant);
    memcpy(m.data(), delays, sizeof(double) * 3 * c->p.nant);
    memcpy(m.data() + 3 * c->p.nant, fileoffsets, sizeof(double) * c->p.nant);
    HIPCHK(hipMemcpyAsync(c->d_model, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_upload_data(fx_context* c, const uint8_t* const* data, long long nbytes)
{
    if (!c || !data) return fail(FX_EINVAL, "null argument");
    if (nbytes < 0 || nbytes > c->cap_bytes)
        return fail(FX_EINVAL, "nbytes=%lld exceeds capacity %lld", nbytes, c->cap_bytes);
    for (int a = 0; a < c->p.nant; a++) {
        if (!data[a]) return fail(FX_EINVAL, "null data for antenna %d", a);
        HIPCHK(hipMemcpyAsync(c->d_data + (size_t)a * c->stride, data[a], (size_t)nbytes, hipMemcpyHostToDevice,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_upload_data_strided(fx_context* c, const uint8_t* data, long long stride, long long nbytes)
{
    if (!c || !data) return fail(FX_EINVAL, "null argument");
    if (nbytes < 0 || nbytes > c->cap_bytes || stride < nbytes)
        return fail(FX_EINVAL, "nbytes=%lld stride=%lld capacity %lld", nbytes, stride, c->cap_bytes);
    HIPCHK(hipMemcpy2DAsync(c->d_data, (size_t)c->stride, data, (size_t)stride, (size_t)nbytes, (size_t)c->p.nant,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_device_data(fx_context* c, void** dptr, long long* stride)
{
    if (!c || !dptr || !stride) return fail(FX_EINVAL, "null argument");
    *dptr = c->d_data;
    *stride = c->stride;
    return FX_OK;
}

static int ev_begin(fx_context* c, int kind, hipStream_t s, EvPair** out)
{
    *out = nullptr;
    if (!c->timing || c->capturing) return FX_OK;
    EvPair e;
    if (!c->pool.empty()) {
        e = c->pool.back();
        c->pool.pop_back();
    } else {
        HIPCHK(hipEventCreate(&e.a));
        HIPCHK(hipEventCreate(&e.b));
    }
    e.kind = kind;
    HIPCHK(hipEventRecord(e.a, s));
    c->pending.push_back(e);
    *out = &c->pending.back();
    return FX_OK;
}

static int ev_end(fx_context* c, hipStream_t s)
{
    if (!c->timing || c->capturing) return FX_OK;
    HIPCHK(hipEventRecord(c->pending.back().b, s));
    return FX_OK;
}

static int ev_collect(fx_context* c)
{
    if (c->pending.empty()) return FX_OK;
    for (EvPair& e : c->pending) {
        float ms = 0.0f;
        HIPCHK(hipEventSynchronize(e.b));
        HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
        c->ms[e.kind] += ms;
        c->pool.push_back(e);
    }
    c->pending.clear();
    return FX_OK;
}

#define TIMED(kind, s, launch)                   \
    do {                                         \
        EvPair* ep_;                             \
        int rc_ = ev_begin(c, (kind), (s), &ep_); \
        if (rc_) return rc_;                     \
        HIPCHK(launch);                          \
        rc_ = ev_end(c, (s));                    \
        if (rc_) return rc_;                     \
    } while (0)

// Host-side checks of one F-stage call; allocates the tap buffer when needed
// (never inside a graph capture).
struct FStageCall {
    int numffts, npad, maxoffset, fft0, origin, slot;
    int contig, it0;    // iteration[k] = it0 + k: nothing to upload
};

static int fstage_check(fx_context* c, int numffts, const int* iteration, int split, int fft0, FStageCall* f,
                        int origin = 0, int slot = 0)
{
    if (!c || !iteration) return fail(FX_EINVAL, "null argument");
    if (numffts < 1 || split < 1) return fail(FX_EINVAL, "numffts=%d split=%d", numffts, split);
    if (fft0 < 0 || fft0 % c->granule) return fail(FX_EINVAL, "fft0=%d not a multiple of %d", fft0, c->granule);
    const int npad = (numffts + c->granule - 1) / c->granule * c->granule;
    if (fft0 + npad > c->max_ffts)
        return fail(FX_EINVAL, "FFT slots [%d,%d) exceed capacity %d", fft0, fft0 + npad, c->max_ffts);
    const FxPlan& p = c->p;
    const long long maxoffset = ((long long)numffts * split - 1) * p.F;
    if (maxoffset > 0x7fffffffLL - 2LL * p.F) return fail(FX_EINVAL, "subint too long for 32-bit offsets");
    // The kernel reads samples [offset, offset + F) with offset <= maxoffset.
    const long long need = p.iscomplex ? 4LL * numffts * split * p.F : ((long long)numffts * split * p.F + 1) / 2;
    if (need > c->cap_bytes)
        return fail(FX_EINVAL, "input needs %lld bytes per antenna, capacity %lld", need, c->cap_bytes);
    if (slot < 0 || slot > 1 || (slot == 1 && !c->d_data1))
        return fail(FX_EINVAL, "input slot %d not set up (fx_ingest_host first)", slot);
    for (int k = 0; k < numffts; k++) {
        const long long rel = ((long long)iteration[k] - origin) * p.F;
        if (iteration[k] < 0 || rel > 0x7fffffffLL - p.F || rel < -0x7fffffffLL + p.F)
            return fail(FX_EINVAL, "iteration[%d]=%d out of range (origin %d)", k, iteration[k], origin);
    }
    if (c->tap && !c->d_tap) {
        if (hipMalloc((void**)&c->d_tap, sizeof(float2) * (size_t)c->max_ffts * p.nant * 2 * p.nchan) != hipSuccess)
            return fail(FX_ENOMEM, "tap allocation failed");
    }
    if ((c->tap & FX_TAP_LEVELS) && !c->d_ltap) {
        if (hipMalloc((void**)&c->d_ltap, sizeof(float) * (size_t)c->max_ffts * p.nant * 2 * p.F) != hipSuccess)
            return fail(FX_ENOMEM, "level tap allocation failed");
    }
    f->numffts = numffts;
    f->npad = npad;
    f->maxoffset = (int)maxoffset;
    f->fft0 = fft0;
    f->origin = origin;
    f->slot = slot;
    return FX_OK;
}

// FFT indices and buffer origin for the delay kernel.  A contiguous run of
// indices (the usual case) travels as kernel arguments.  A launch recorded
// into a graph must read each subint's values from device memory (device =
// true): the block {origin, first index, list flag} at d_iter + max_ffts is
// copied from a pinned ring slot (an asynchronous DMA, no host staging), and
// only a non-contiguous index list is copied to d_iter itself.
static int iter_upload(fx_context* c, const int* iteration, FStageCall& f, hipStream_t s, bool device = false)
{
    f.it0 = iteration[0];
    bool contig = true;
    for (int k = 1; k < f.numffts && contig; k++) contig = iteration[k] == f.it0 + k;
    f.contig = contig && !device;
    if (f.contig) return FX_OK;
    if (!contig)
        HIPCHK(hipMemcpyAsync(c->d_iter, iteration, sizeof(int) * f.numffts, hipMemcpyHostToDevice, s));
    const int slot = c->ring_pos;
    c->ring_pos = (c->ring_pos + 1) % kRing;
    if (c->ring_used[slot]) HIPCHK(hipEventSynchronize(c->ring_ev[slot]));   // its last copy has been made
    int* blk = c->h_ring + 4 * slot;
    blk[0] = f.origin;
    blk[1] = f.it0;
    blk[2] = contig ? 0 : 1;
    blk[3] = 0;
    HIPCHK(hipMemcpyAsync(c->d_iter + c->max_ffts, blk, sizeof(int) * 3, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(c->ring_ev[slot], s));
    c->ring_used[slot] = 1;
    return FX_OK;
}

static int fstage_enqueue(fx_context* c, const FStageCall& f, hipStream_t s)
{
    const FxPlan& p = c->p;
    // the slot's input has arrived (graphs of batches record kernels only and
    // read slot 0, filled by plain uploads)
    const bool ingest = c->ingest && !c->capturing;
    if (ingest) HIPCHK(hipStreamWaitEvent(s, c->copied[f.slot], 0));
    TIMED(TM_DELAY, s, launch_delay_rotators(p, c->d_model, c->d_model + 3 * p.nant, f.contig ? nullptr : c->d_iter,
                                             f.it0, f.numffts, f.npad, f.maxoffset,
                                             f.contig ? nullptr : c->d_iter + c->max_ffts, f.origin, c->d_subchf,
                                             c->d_stepchf, fstage_step_q(p.lg, c->fplan, p.iscomplex), c->d_dtab,
                                             c->d_rtab, c->d_valid + (size_t)f.fft0 * p.nant, s));
    FStageArgs a;
    a.p = p;
    a.data = f.slot ? c->d_data1 : c->d_data;
    a.data_stride = c->stride;
    a.delays = c->d_dtab;
    a.rtab = c->d_rtab;
    a.tw = c->d_tw;
    a.subchf = c->d_subchf;
    a.stepchf = c->d_stepchf;
    a.out = c->d_xin;
    a.kc = c->kgroups;
    a.fg0 = f.fft0 / c->G;
    a.ngroups = f.npad / c->G;
    a.scale_inv = c->d_scale_inv;
    a.tap = c->tap ? c->d_tap : nullptr;
    a.ltap = (c->tap & FX_TAP_LEVELS) && !p.iscomplex ? c->d_ltap : nullptr;
    a.stage = c->d_stage;
    a.plan = c->fplan;
    TIMED(TM_FSTAGE, s, launch_fstage(a, c->fprec, s));
    if (ingest) HIPCHK(hipEventRecord(c->slot_free[f.slot], s));           // the slot may be overwritten
    if (c->timing && !c->capturing) c->nl[0]++;
    c->last_nffts = f.npad;
    return FX_OK;
}

extern "C" int fx_fstage(fx_context* c, int numffts, const int* iteration, int split, int fft0, int* nslots)
{
    FStageCall f;
    int rc = fstage_check(c, numffts, iteration, split, fft0, &f);
    if (rc) return rc;
    rc = iter_upload(c, iteration, f, c->stream);
    if (rc) return rc;
    rc = fstage_enqueue(c, f, c->stream);
    if (rc) return rc;
    if (nslots) *nslots = f.npad;
    return FX_OK;
}

extern "C" int fx_fstage_ex(fx_context* c, int slot, int numffts, const int* iteration, int split, int origin,
                            int fft0, int* nslots)
{
    FStageCall f;
    int rc = fstage_check(c, numffts, iteration, split, fft0, &f, origin, slot);
    if (rc) return rc;
    rc = iter_upload(c, iteration, f, c->stream);
    if (rc) return rc;
    rc = fstage_enqueue(c, f, c->stream);
    if (rc) return rc;
    if (nslots) *nslots = f.npad;
    return FX_OK;
}

// Streaming ingest: two input slots, each a device buffer and a pinned host
// staging buffer; H2D copies run on their own stream and overlap the
// F-stage / X-engine work on the context's stream.
static int ingest_setup(fx_context* c)
{
    if (c->ingest) return FX_OK;
    const size_t bytes = (size_t)c->stride * c->p.nant;
    if (!c->d_data1 && hipMalloc((void**)&c->d_data1, bytes) != hipSuccess)
        return fail(FX_ENOMEM, "second input slot (%zu bytes) failed", bytes);
    for (int k = 0; k < 2; k++)
        if (!c->h_pin[k] && hipHostMalloc((void**)&c->h_pin[k], bytes, hipHostMallocDefault) != hipSuccess)
            return fail(FX_ENOMEM, "pinned staging buffer (%zu bytes) failed", bytes);
    if (!c->cstream) HIPCHK(hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++) {
        if (!c->copied[k]) HIPCHK(hipEventCreateWithFlags(&c->copied[k], hipEventDisableTiming));
        if (!c->slot_free[k]) HIPCHK(hipEventCreateWithFlags(&c->slot_free[k], hipEventDisableTiming));
        // both slots start free and filled
        HIPCHK(hipEventRecord(c->copied[k], c->cstream));
        HIPCHK(hipEventRecord(c->slot_free[k], c->stream));
    }
    c->ingest = 1;
    return FX_OK;
}

extern "C" int fx_ingest_host(fx_context* c, int slot, uint8_t** ptr, long long* stride)
{
    if (!c || !ptr || !stride) return fail(FX_EINVAL, "null argument");
    if (slot < 0 || slot > 1) return fail(FX_EINVAL, "slot %d (0 or 1)", slot);
    int rc = ingest_setup(c);
    if (rc) return rc;
    *ptr = c->h_pin[slot];
    *stride = c->stride;
    return FX_OK;
}

extern "C" int fx_ingest_submit(fx_context* c, int slot, long long nbytes)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (!c->ingest) return fail(FX_EINVAL, "fx_ingest_host must come first");
    if (slot < 0 || slot > 1) return fail(FX_EINVAL, "slot %d (0 or 1)", slot);
    if (nbytes < 0 || nbytes > c->cap_bytes)
        return fail(FX_EINVAL, "nbytes=%lld exceeds capacity %lld", nbytes, c->cap_bytes);
    uint8_t* dst = slot ? c->d_data1 : c->d_data;
    HIPCHK(hipStreamWaitEvent(c->cstream, c->slot_free[slot], 0));       // last F-stage on the slot is done
    HIPCHK(hipMemcpy2DAsync(dst, (size_t)c->stride, c->h_pin[slot], (size_t)c->stride, (size_t)nbytes,
                            (size_t)c->p.nant, hipMemcpyHostToDevice, c->cstream));
    HIPCHK(hipEventRecord(c->copied[slot], c->cstream));
    return FX_OK;
}

extern "C" int fx_ingest_wait(fx_context* c, int slot)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (!c->ingest) return FX_OK;
    if (slot < 0 || slot > 1) return fail(FX_EINVAL, "slot %d (0 or 1)", slot);
    HIPCHK(hipEventSynchronize(c->copied[slot]));
    return FX_OK;
}

static int xstage_check(fx_context* c, int fft0, int nslots)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (fft0 < 0 || nslots < 0 || fft0 % c->granule || nslots % c->granule || fft0 + nslots > c->max_ffts)
        return fail(FX_EINVAL, "bad slot range [%d,%d) (granule %d, capacity %d)", fft0, fft0 + nslots, c->granule,
                    c->max_ffts);
    return FX_OK;
}

static int xstage_enqueue(fx_context* c, int fft0, int nslots, hipStream_t s)
{
    if (nslots == 0) return FX_OK;
    XArgs a;
    a.p = c->p;
    a.xin = c->d_xin;
    a.fg0 = fft0 / c->G;
    a.ngroups = nslots / c->G;
    a.vacc = c->d_vacc;
    a.accumulate = 1;
    a.scale = c->d_scale;
    a.sb = c->xsb;
    TIMED(TM_XSTAGE, s, launch_xengine(a, c->prec, s));
    // valid-FFT counts of the same slots (kernel.c:850)
    TIMED(TM_COUNTS, s, launch_counts(c->p.nant, nslots, c->d_valid + (size_t)fft0 * c->p.nant, c->d_counts, s));
    // FP8: the next F-stages scale by the autocorrelations accumulated so far
    if (c->autoscale && c->prec == FX_PREC_F8)
        HIPCHK(launch_fp8_autoscale(c->p, c->d_vacc, c->d_counts, c->d_scale, c->d_scale_inv, kFp8Target, s));
    if (c->timing && !c->capturing) c->nl[1]++;
    return FX_OK;
}

extern "C" int fx_xstage(fx_context* c, int fft0, int nslots)
{
    int rc = xstage_check(c, fft0, nslots);
    if (rc) return rc;
    return xstage_enqueue(c, fft0, nslots, c->stream);
}

// One subint as one hipGraph launch (fx_subint).  The graph holds the delay /
// rotator kernel, the F-stage and, on batch boundaries, the X-engine and the
// counts; it is recorded by stream capture on the first call of a shape and
// replayed after that.  A contiguous run of FFT indices (the usual case)
// reaches the delay kernel as two kernel arguments, the first index and the
// buffer origin, which each launch updates in its executable graph
// (hipGraphExecKernelNodeSetParams: no copy, no extra GPU command).  Every
// shape is instantiated kExecCopies times and used round robin, and a copy's
// arguments are updated only after its previous launch has completed, so an
// update never meets a launch still queued.  Other index lists are copied to
// the device ahead of the launch and read there.  Waits on a streaming slot's
// copy and the slot's release stay outside the graph.
static constexpr size_t kMaxSubGraphs = 64;

static int subgraph_record(fx_context* c, const FStageCall& f, int xslots, bool args, fx_context::SubGraph& sg)
{
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = 1;
    int rc = fstage_enqueue(c, f, c->stream);
    if (rc == FX_OK && xslots > 0) rc = xstage_enqueue(c, 0, xslots, c->stream);
    c->capturing = 0;
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(c->stream, &g);
    if (rc != FX_OK || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != FX_OK ? rc : fail(FX_EHIP, "subint graph capture failed: %s", hipGetErrorString(ee));
    }
    sg.graph = g;
    if (args) {
        size_t n = 0;
        HIPCHK(hipGraphGetNodes(g, nullptr, &n));
        std::vector<hipGraphNode_t> nodes(n);
        HIPCHK(hipGraphGetNodes(g, nodes.data(), &n));
        for (hipGraphNode_t nd : nodes) {
            hipGraphNodeType t;
            if (hipGraphNodeGetType(nd, &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
            hipKernelNodeParams kp;
            if (hipGraphKernelNodeGetParams(nd, &kp) != hipSuccess) continue;
            if (kp.func == delay_rot_kernel_fn() && kp.kernelParams) {
                sg.dnode = nd;
                sg.kp = kp;
            }
        }
        if (!sg.dnode) return FX_EUNSUPPORTED;              // the caller records with device indices instead
    }
    for (int i = 0; i < kExecCopies; i++) {
        const hipError_t ie = hipGraphInstantiate(&sg.exec[i], g, nullptr, nullptr, 0);
        if (ie != hipSuccess) return fail(FX_EHIP, "subint graph instantiation failed: %s", hipGetErrorString(ie));
        HIPCHK(hipEventCreateWithFlags(&sg.done[i], hipEventDisableTiming));
    }
    return FX_OK;
}

static int subint_graph(fx_context* c, FStageCall& f, const int* iteration, int xslots)
{
    // contiguous indices as kernel arguments where the runtime lets the
    // recorded node's arguments be updated; otherwise from device memory
    bool args = c->graph_args;
    for (int k = 1; k < f.numffts && args; k++) args = iteration[k] == iteration[0] + k;
    for (int attempt = 0; attempt < 2; attempt++) {
        if (args) {
            f.it0 = iteration[0];
            f.contig = 1;
        } else {
            int rc = iter_upload(c, iteration, f, c->stream, true);
            if (rc) return rc;
        }
        const long long key[10] = {f.numffts, f.maxoffset, f.slot, f.fft0, xslots, c->tap, c->xsb, c->fplan,
                                   c->autoscale, args ? 1 : 0};
        fx_context::SubGraph* sg = nullptr;
        for (auto& g : c->graphs)
            if (std::equal(key, key + 10, g.key)) sg = &g;
        if (!sg) {
            if (c->graphs.size() >= kMaxSubGraphs) {
                HIPCHK(hipStreamSynchronize(c->stream));
                for (auto& g : c->graphs) subgraph_free(g);
                c->graphs.clear();
            }
            fx_context::SubGraph ng{};
            std::copy(key, key + 10, ng.key);
            const int rc = subgraph_record(c, f, xslots, args, ng);
            if (rc == FX_EUNSUPPORTED && args) {
                // no updatable argument list from this runtime: device indices from now on
                subgraph_free(ng);
                c->graph_args = 0;
                args = false;
                continue;
            }
            if (rc) {
                subgraph_free(ng);
                return rc;
            }
            c->graphs.push_back(ng);
            sg = &c->graphs.back();
        }
        const int i = sg->next;
        sg->next = (sg->next + 1) % kExecCopies;
        if (sg->used[i]) HIPCHK(hipEventSynchronize(sg->done[i]));   // this copy's last launch is done
        if (args) {
            int it0 = f.it0, org = f.origin;
            void* a[kDelayNArgs];
            for (int k = 0; k < kDelayNArgs; k++) a[k] = sg->kp.kernelParams[k];
            a[kDelayArgIt0] = &it0;
            a[kDelayArgOrigin] = &org;
            hipKernelNodeParams kp = sg->kp;
            kp.kernelParams = a;
            kp.extra = nullptr;
            HIPCHK(hipGraphExecKernelNodeSetParams(sg->exec[i], sg->dnode, &kp));
        }
        const bool ingest = c->ingest;
        if (ingest) HIPCHK(hipStreamWaitEvent(c->stream, c->copied[f.slot], 0));
        TIMED(TM_GRAPH, c->stream, hipGraphLaunch(sg->exec[i], c->stream));
        HIPCHK(hipEventRecord(sg->done[i], c->stream));
        sg->used[i] = 1;
        if (ingest) HIPCHK(hipEventRecord(c->slot_free[f.slot], c->stream));
        if (c->timing) c->nl[2]++;
        c->last_nffts = f.npad;
        return FX_OK;
    }
    return fail(FX_EHIP, "subint graph: no recording");
}

extern "C" int fx_subint(fx_context* c, int slot, int numffts, const int* iteration, int split, int origin, int fft0,
                         int xslots, int* nslots)
{
    FStageCall f;
    int rc = fstage_check(c, numffts, iteration, split, fft0, &f, origin, slot);
    if (rc) return rc;
    if (xslots < 0 || (xslots > 0 && (xslots % c->granule || xslots > c->max_ffts || xslots < fft0 + f.npad)))
        return fail(FX_EINVAL, "xslots=%d (granule %d, capacity %d, must cover the subint's slots [%d,%d))", xslots,
                    c->granule, c->max_ffts, fft0, fft0 + f.npad);
    if (nslots) *nslots = f.npad;
    if (!c->use_graph) {
        rc = iter_upload(c, iteration, f, c->stream);
        if (rc == FX_OK) rc = fstage_enqueue(c, f, c->stream);
        if (rc == FX_OK && xslots > 0) rc = xstage_enqueue(c, 0, xslots, c->stream);
        return rc;
    }
    return subint_graph(c, f, iteration, xslots);
}

extern "C" int fx_process(fx_context* c, int numffts, const int* iteration, int split)
{
    if (!c || !iteration) return fail(FX_EINVAL, "null argument");
    if (numffts < 1) return fail(FX_EINVAL, "numffts=%d", numffts);
    const int npad = (numffts + c->granule - 1) / c->granule * c->granule;
    return fx_subint(c, 0, numffts, iteration, split, 0, 0, npad, nullptr);
}

extern "C" int fx_reset(fx_context* c)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    HIPCHK(launch_zero(c->d_vacc, c->vacc_bytes, c->stream));
    HIPCHK(launch_zero(c->d_counts, sizeof(int) * c->p.nant * c->p.nant, c->stream));
    return FX_OK;
}

extern "C" int fx_read_vis_ex(fx_context* c, int flags, float* cross, int* cross_counts, float* autos,
                              int* auto_counts)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (flags & ~FX_VIS_NORMALIZE) return fail(FX_EINVAL, "unknown flags 0x%x", flags);
    const FxPlan& p = c->p;
    const int nbl = p.nant * (p.nant - 1) / 2;
    HIPCHK(launch_export(p, c->d_vacc, c->d_cross, c->d_autos, (flags & FX_VIS_NORMALIZE) ? c->d_counts : nullptr,
                         c->stream));
    if (cross && nbl)
        HIPCHK(hipMemcpyAsync(cross, c->d_cross, sizeof(float2) * (size_t)nbl * 4 * p.nchan, hipMemcpyDeviceToHost,
                              c->stream));
    if (autos)
        HIPCHK(hipMemcpyAsync(autos, c->d_autos, sizeof(float2) * (size_t)p.nant * 4 * p.nchan,
                              hipMemcpyDeviceToHost, c->stream));
    std::vector<int> cnt((size_t)p.nant * p.nant);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->d_counts, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (cross_counts) {
        int b = 0;
        for (int j = 0; j < p.nant; j++)
            for (int k = j + 1; k < p.nant; k++) cross_counts[b++] = cnt[(size_t)j * p.nant + k];
    }
    if (auto_counts)
        for (int j = 0; j < p.nant; j++) auto_counts[j] = cnt[(size_t)j * p.nant + j];
    return FX_OK;
}

extern "C" int fx_read_vis(fx_context* c, float* cross, int* cross_counts, float* autos, int* auto_counts)
{
    return fx_read_vis_ex(c, 0, cross, cross_counts, autos, auto_counts);
}

// Visibility files.  Binary layout (little endian):
//   char magic[8] = "DIFXFXV1"; int32 nant, nchan, nbl, flags; double bw, lo;
//   int32 cross_counts[nbl]; int32 auto_counts[nant];
//   float cross[nbl][4][nchan][2]; float autos[nant][4][nchan][2]
// Text: the reference's vis.out (kernel.c:608-617) of the cross visibilities,
// one line per channel, lines formatted in parallel.
static int write_text(const char* path, const FxPlan& p, const std::vector<float2>& cross)
{
    const int nbl = p.nant * (p.nant - 1) / 2;
    FILE* f = fopen(path, "w");
    if (!f) return fail(FX_EINVAL, "cannot open %s", path);
    const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int chunk = 64;                                  // channels formatted per round and thread
    std::vector<std::string> lines((size_t)nth * chunk);
    int rc = FX_OK;
    for (int c0 = 0; c0 < p.nchan && rc == FX_OK; c0 += nth * chunk) {
        const int c1 = std::min(p.nchan, c0 + nth * chunk);
        auto work = [&](int t) {
            char num[64];
            for (int c = c0 + t; c < c1; c += nth) {
                std::string& L = lines[c - c0];
                L.clear();
                snprintf(num, sizeof(num), "%5d %11.6f", c, (c + 0.5) / p.nchan * p.bw / 1e6);
                L += num;
                for (int b = 0; b < nbl; b++)
                    for (int q = 0; q < 4; q++) {
                        const float2 v = cross[((size_t)b * 4 + q) * p.nchan + c];
                        const float amp = (float)sqrt((double)v.x * v.x + (double)v.y * v.y);
                        const float ph = (float)atan2((double)v.y, (double)v.x);
                        snprintf(num, sizeof(num), " %11.6f %11.6f %11.6f %10.6f", v.x, v.y, amp, ph);
                        L += num;
                    }
                L += '\n';
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nth; t++) th.emplace_back(work, t);
        work(0);
        for (std::thread& x : th) x.join();
        for (int c = c0; c < c1; c++)
            if (fwrite(lines[c - c0].data(), 1, lines[c - c0].size(), f) != lines[c - c0].size()) rc = FX_EINVAL;
    }
    if (fclose(f) != 0 || rc) return fail(FX_EINVAL, "writing %s failed", path);
    return FX_OK;
}

extern "C" int fx_save_vis(fx_context* c, const char* path, int format, int flags)
{
    if (!c || !path) return fail(FX_EINVAL, "null argument");
    if (format != FX_VIS_BINARY && format != FX_VIS_TEXT) return fail(FX_EINVAL, "unknown format %d", format);
    const FxPlan& p = c->p;
    const int nbl = p.nant * (p.nant - 1) / 2;
    std::vector<float2> cross((size_t)std::max(nbl, 1) * 4 * p.nchan), autos((size_t)p.nant * 4 * p.nchan);
    std::vector<int> cc(std::max(nbl, 1)), ac(p.nant);
    int rc = fx_read_vis_ex(c, flags, reinterpret_cast<float*>(cross.data()), cc.data(),
                            reinterpret_cast<float*>(autos.data()), ac.data());
    if (rc) return rc;
    if (format == FX_VIS_TEXT) return write_text(path, p, cross);
    FILE* f = fopen(path, "wb");
    if (!f) return fail(FX_EINVAL, "cannot open %s", path);
    const int32_t hdr[4] = {p.nant, p.nchan, nbl, flags};
    const double fb[2] = {p.bw, p.lo};
    bool ok = fwrite("DIFXFXV1", 1, 8, f) == 8 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
              fwrite(fb, sizeof(fb), 1, f) == 1 && fwrite(cc.data(), sizeof(int), nbl, f) == (size_t)nbl &&
              fwrite(ac.data(), sizeof(int), p.nant, f) == (size_t)p.nant &&
              fwrite(cross.data(), sizeof(float2), (size_t)nbl * 4 * p.nchan, f) == (size_t)nbl * 4 * p.nchan &&
              fwrite(autos.data(), sizeof(float2), autos.size(), f) == autos.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? FX_OK : fail(FX_EINVAL, "writing %s failed", path);
}

extern "C" int fx_read_delay_table(fx_context* c, int nffts, void* out)
{
    if (!c || !out) return fail(FX_EINVAL, "null argument");
    if (nffts < 0 || nffts > c->max_ffts) return fail(FX_EINVAL, "nffts=%d", nffts);
    HIPCHK(hipMemcpyAsync(out, c->d_dtab, sizeof(FxDelay) * (size_t)nffts * c->p.nant, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_enable_tap(fx_context* c, int enable)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (enable & ~(FX_TAP_SPECTRA | FX_TAP_LEVELS)) return fail(FX_EINVAL, "unknown tap bits 0x%x", enable);
    // the levels ride on the spectra tap's kernel variant
    c->tap = enable ? (enable | FX_TAP_SPECTRA) : 0;
    return FX_OK;
}

extern "C" int fx_read_tap(fx_context* c, int nffts, float* out)
{
    if (!c || !out) return fail(FX_EINVAL, "null argument");
    if (!c->d_tap) return fail(FX_EINVAL, "tap not enabled before fx_fstage");
    if (nffts < 0 || nffts > c->max_ffts) return fail(FX_EINVAL, "nffts=%d", nffts);
    HIPCHK(hipMemcpyAsync(out, c->d_tap, sizeof(float2) * (size_t)nffts * c->p.nant * 2 * c->p.nchan,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_read_tap_levels(fx_context* c, int nffts, float* out)
{
    if (!c || !out) return fail(FX_EINVAL, "null argument");
    if (c->p.iscomplex) return fail(FX_EUNSUPPORTED, "level tap: real 2-bit input only");
    if (!c->d_ltap) return fail(FX_EINVAL, "level tap not enabled before fx_fstage");
    if (nffts < 0 || nffts > c->max_ffts) return fail(FX_EINVAL, "nffts=%d", nffts);
    HIPCHK(hipMemcpyAsync(out, c->d_ltap, sizeof(float) * (size_t)nffts * c->p.nant * 2 * c->p.F,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_set_fp8_scales(fx_context* c, const float* scales)
{
    if (!c || !scales) return fail(FX_EINVAL, "null argument");
    const FxPlan& p = c->p;
    std::vector<float> sc((size_t)p.nchan * p.Spad, 1.0f), sci((size_t)p.nchan * p.Spad, 1.0f);
    for (int ch = 0; ch < p.nchan; ch++)
        for (int s = 0; s < p.S; s++) {
            const float v = scales[(size_t)ch * p.S + s];
            if (!(v > 0.0f) || !std::isfinite(v)) return fail(FX_EINVAL, "scale[%d][%d] not positive", ch, s);
            sc[(size_t)ch * p.Spad + s] = v;
            sci[(size_t)s * p.nchan + ch] = 1.0f / v;
        }
    HIPCHK(hipMemcpyAsync(c->d_scale, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_scale_inv, sci.data(), sizeof(float) * sci.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_update_fp8_scales(fx_context* c)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (c->prec != FX_PREC_F8) return fail(FX_EINVAL, "FP8 scales: context precision is not FX_PREC_F8");
    HIPCHK(launch_fp8_autoscale(c->p, c->d_vacc, c->d_counts, c->d_scale, c->d_scale_inv, kFp8Target, c->stream));
    return FX_OK;
}

extern "C" int fx_sync(fx_context* c)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    return FX_OK;
}

extern "C" int fx_set_timing(fx_context* c, int enable)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    c->timing = enable ? 1 : 0;
    return FX_OK;
}

extern "C" int fx_reset_timing(fx_context* c)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    int rc = ev_collect(c);
    if (rc) return rc;
    for (double& v : c->ms) v = 0.0;
    c->nl[0] = c->nl[1] = c->nl[2] = 0;
    return FX_OK;
}

extern "C" int fx_get_timing(fx_context* c, double* ms4, int* nlaunch2)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    int rc = ev_collect(c);
    if (rc) return rc;
    if (ms4)
        for (int i = 0; i < 4; i++) ms4[i] = c->ms[i];
    if (nlaunch2) {
        nlaunch2[0] = c->nl[0];
        nlaunch2[1] = c->nl[1];
    }
    return FX_OK;
}

extern "C" int fx_get_graph_timing(fx_context* c, double* ms, int* nlaunch)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    int rc = ev_collect(c);
    if (rc) return rc;
    if (ms) *ms = c->ms[TM_GRAPH];
    if (nlaunch) *nlaunch = c->nl[2];
    return FX_OK;
}

extern "C" int fx_get_geometry(const fx_context* c, int* o)
{
    if (!c || !o) return fail(FX_EINVAL, "null argument");
    const FxPlan& p = c->p;
    const int v[11] = {p.F, p.stride, p.sub, p.nstep, p.nfstep, p.S, p.Spad, p.nblk, p.nbp, c->G, c->granule};
    memcpy(o, v, sizeof(v));
    return FX_OK;
}

// --------------------------------------------------------------------------
// Tuning options

extern "C" int fx_set_option(fx_context* c, int option, int value)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    switch (option) {
    case FX_OPT_XTILE:
        if (value != 0 && value != 1 && value != 2 && value != 4)
            return fail(FX_EINVAL, "X-engine tile %d: expected 0 (automatic), 1, 2 or 4", value);
        c->xsb = value;
        return FX_OK;
    case FX_OPT_FPLAN:
        if (value < 0 || value >= fstage_plans(c->p.lg, c->p.iscomplex))
            return fail(FX_EINVAL, "F-stage plan %d: %d plans at F=%d", value, fstage_plans(c->p.lg, c->p.iscomplex), c->p.F);
        c->fplan = value;
        return FX_OK;
    case FX_OPT_XBATCH:
        if (value < 1) return fail(FX_EINVAL, "subints per X-engine launch must be >= 1");
        c->xbatch = value;
        return FX_OK;
    case FX_OPT_GRAPH:
        if (value != 0 && value != 1) return fail(FX_EINVAL, "FX_OPT_GRAPH is 0 or 1");
        c->use_graph = value;
        return FX_OK;
    case FX_OPT_FP8_AUTOSCALE:
        if (value != 0 && value != 1) return fail(FX_EINVAL, "FX_OPT_FP8_AUTOSCALE is 0 or 1");
        if (value && c->prec != FX_PREC_F8) return fail(FX_EINVAL, "FX_OPT_FP8_AUTOSCALE needs FX_PREC_F8");
        c->autoscale = value;
        return FX_OK;
    default: return fail(FX_EINVAL, "unknown option %d", option);
    }
}

extern "C" int fx_get_option(const fx_context* c, int option, int* value)
{
    if (!c || !value) return fail(FX_EINVAL, "null argument");
    switch (option) {
    case FX_OPT_XTILE: *value = c->xsb > 0 ? c->xsb : xengine_super_blocks(c->p.nblk); return FX_OK;
    case FX_OPT_FPLAN: *value = c->fplan; return FX_OK;
    case FX_OPT_NPLANS: *value = fstage_plans(c->p.lg, c->p.iscomplex); return FX_OK;
    case FX_OPT_XBATCH: *value = c->xbatch; return FX_OK;
    case FX_OPT_GRAPH: *value = c->use_graph; return FX_OK;
    case FX_OPT_FP8_AUTOSCALE: *value = c->autoscale; return FX_OK;
    default: return fail(FX_EINVAL, "unknown option %d", option);
    }
}

// --------------------------------------------------------------------------
// Batches of contexts (one per subband): one subint of every subband per
// call, as one hipGraph launch whose branches are the subbands, or as plain
// launches one subband after another.

struct fx_batch {
    std::vector<fx_context*> ctx;
    hipStream_t s0 = nullptr;                 // launch stream
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> pre, join;        // per context
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<long long> key;               // call shape the graph was captured for
    int captures = 0;
};

static void batch_drop_graph(fx_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key.clear();
}

static void batch_free(fx_batch* b)
{
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    batch_drop_graph(b);
    for (hipEvent_t e : b->pre) if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->join) if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    delete b;
}

extern "C" int fx_batch_create(fx_context* const* ctxs, int n, fx_batch** out)
{
    if (!ctxs || !out || n < 1) return fail(FX_EINVAL, "need at least one context");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return fail(FX_EINVAL, "context %d is null", i);
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return fail(FX_EINVAL, "context %d appears twice", i);
    }
    fx_batch* b = new fx_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->pre.assign(n, nullptr);
    b->join.assign(n, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < n; i++)
        ok = hipEventCreateWithFlags(&b->pre[i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        batch_free(b);
        return fail(FX_EHIP, "batch stream/event creation failed");
    }
    *out = b;
    return FX_OK;
}

extern "C" void fx_batch_destroy(fx_batch* b)
{
    if (b) batch_free(b);
}

// Records the subint of every context into a graph on b->s0 (subbands as
// branches forked from and joined back to s0).
static int batch_capture(fx_batch* b, const std::vector<FStageCall>& calls)
{
    const int n = (int)b->ctx.size();
    batch_drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal));
    int rc = FX_OK;
    hipError_t he = hipEventRecord(b->fork, b->s0);
    for (int i = 0; he == hipSuccess && rc == FX_OK && i < n; i++) {
        fx_context* c = b->ctx[i];
        he = hipStreamWaitEvent(c->stream, b->fork, 0);
        if (he != hipSuccess) break;
        c->capturing = 1;
        rc = fstage_enqueue(c, calls[i], c->stream);
        if (rc == FX_OK) rc = xstage_enqueue(c, 0, calls[i].npad, c->stream);
        c->capturing = 0;
        if (rc == FX_OK) he = hipEventRecord(b->join[i], c->stream);
        if (rc == FX_OK && he == hipSuccess) he = hipStreamWaitEvent(b->s0, b->join[i], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->s0, &g);
    if (rc != FX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(FX_EHIP, "graph capture failed: %s", hipGetErrorString(he != hipSuccess ? he : ee));
    }
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->captures++;
    return FX_OK;
}

extern "C" int fx_batch_process_ex(fx_batch* b, const int* numffts, const int* const* iteration,
                                   const int* origins, int split, int mode)
{
    if (!b || !numffts || !iteration) return fail(FX_EINVAL, "null argument");
    if (mode != FX_BATCH_GRAPH && mode != FX_BATCH_LAUNCH) return fail(FX_EINVAL, "bad mode %d", mode);
    const int n = (int)b->ctx.size();
    std::vector<FStageCall> calls(n);
    std::vector<long long> key;
    key.reserve(2 * n + 1);
    for (int i = 0; i < n; i++) {
        int rc = fstage_check(b->ctx[i], numffts[i], iteration[i], split, 0, &calls[i], origins ? origins[i] : 0);
        if (rc) return fail(rc, "subband %d: %s", i, fx_last_error());
        key.push_back(numffts[i]);
        key.push_back((((long long)b->ctx[i]->autoscale * 8 + b->ctx[i]->tap) * 8 + b->ctx[i]->xsb) * 8 +
                      b->ctx[i]->fplan);
    }
    key.push_back(split);
    // s0 starts after the work already queued on every context's stream
    for (int i = 0; i < n; i++) {
        HIPCHK(hipEventRecord(b->pre[i], b->ctx[i]->stream));
        HIPCHK(hipStreamWaitEvent(b->s0, b->pre[i], 0));
    }
    for (int i = 0; i < n; i++) {
        int rc = iter_upload(b->ctx[i], iteration[i], calls[i], b->s0, mode == FX_BATCH_GRAPH);
        if (rc) return rc;
    }
    if (mode == FX_BATCH_GRAPH) {
        if (!b->exec || key != b->key) {
            int rc = batch_capture(b, calls);
            if (rc) return rc;
            b->key = key;
        }
        HIPCHK(hipGraphLaunch(b->exec, b->s0));
    } else {
        for (int i = 0; i < n; i++) {
            int rc = fstage_enqueue(b->ctx[i], calls[i], b->s0);
            if (rc == FX_OK) rc = xstage_enqueue(b->ctx[i], 0, calls[i].npad, b->s0);
            if (rc) return rc;
        }
    }
    // later calls on each context's own stream follow this subint
    HIPCHK(hipEventRecord(b->done, b->s0));
    for (int i = 0; i < n; i++) {
        HIPCHK(hipStreamWaitEvent(b->ctx[i]->stream, b->done, 0));
        b->ctx[i]->last_nffts = calls[i].npad;
    }
    return FX_OK;
}

extern "C" int fx_batch_process(fx_batch* b, const int* numffts, const int* const* iteration, int split, int mode)
{
    return fx_batch_process_ex(b, numffts, iteration, nullptr, split, mode);
}

extern "C" int fx_batch_sync(fx_batch* b)
{
    if (!b) return fail(FX_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(b->s0));
    return FX_OK;
}

extern "C" int fx_batch_captures(const fx_batch* b) { return b ? b->captures : FX_EINVAL; }

// --------------------------------------------------------------------------
// Autotuner.  Candidates: every F-stage radix plan, every X-engine tile that
// fits the signal count, and 1, 2, 4, ... subints per X-engine launch up to
// the context's FFT capacity.  Each candidate runs once untimed and `reps`
// times between HIP events on the context's stream.  The cache is a text
// file, one line per (arch, nant, nchan, prec):
//   arch nant nchan prec plan xtile xbatch f_ms x_ms_per_subint

struct TuneEntry {
    std::string arch;
    int nant, nchan, prec, plan, xtile, xbatch;
    double fms, xms;
};

static std::string device_arch()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

static std::vector<TuneEntry> tune_read(const char* path)
{
    std::vector<TuneEntry> v;
    FILE* f = path ? fopen(path, "r") : nullptr;
    if (!f) return v;
    char line[512], arch[128];
    while (fgets(line, sizeof(line), f)) {
        TuneEntry e;
        if (line[0] == '#') continue;
        if (sscanf(line, "%127s %d %d %d %d %d %d %lf %lf", arch, &e.nant, &e.nchan, &e.prec, &e.plan, &e.xtile,
                   &e.xbatch, &e.fms, &e.xms) == 9) {
            e.arch = arch;
            v.push_back(e);
        }
    }
    fclose(f);
    return v;
}

static int tune_write(const char* path, const TuneEntry& e)
{
    std::vector<TuneEntry> v = tune_read(path);
    bool found = false;
    for (TuneEntry& x : v)
        if (x.arch == e.arch && x.nant == e.nant && x.nchan == e.nchan && x.prec == e.prec) {
            x = e;
            found = true;
        }
    if (!found) v.push_back(e);
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return fail(FX_EINVAL, "cannot write tuning cache %s", tmp.c_str());
    fprintf(f, "# arch nant nchan prec plan xtile xbatch f_ms x_ms_per_subint\n");
    for (const TuneEntry& x : v)
        fprintf(f, "%s %d %d %d %d %d %d %.6f %.6f\n", x.arch.c_str(), x.nant, x.nchan, x.prec, x.plan, x.xtile,
                x.xbatch, x.fms, x.xms);
    const bool ok = fclose(f) == 0 && rename(tmp.c_str(), path) == 0;
    return ok ? FX_OK : fail(FX_EINVAL, "cannot write tuning cache %s", path);
}

static void tune_apply_cached(fx_context* c)
{
    const char* path = getenv("FX_TUNE_CACHE");
    if (!path || !*path) return;
    const std::string arch = device_arch();
    for (const TuneEntry& e : tune_read(path))
        if (e.arch == arch && e.nant == c->p.nant && e.nchan == c->p.nchan && e.prec == c->prec) {
            if (e.plan >= 0 && e.plan < fstage_plans(c->p.lg, c->p.iscomplex)) c->fplan = e.plan;
            if (e.xtile == 1 || e.xtile == 2 || e.xtile == 4) c->xsb = e.xtile;
            if (e.xbatch >= 1) c->xbatch = e.xbatch;
        }
}

// Average ms of `reps` runs of fn (after one untimed run) on c->stream.
template <class Fn>
static int tune_time(fx_context* c, int reps, double* ms, Fn fn)
{
    int rc = fn();
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(FX_EHIP, "hipEventCreate failed");
    }
    hipError_t he = hipEventRecord(a, c->stream);
    for (int r = 0; rc == FX_OK && he == hipSuccess && r < reps; r++) rc = fn();
    if (he == hipSuccess) he = hipEventRecord(b, c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(b);
    float t = 0.0f;
    if (he == hipSuccess) he = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (he != hipSuccess) return fail(FX_EHIP, "tuning timer: %s", hipGetErrorString(he));
    *ms = t / reps;
    return FX_OK;
}

static int autotune_run(fx_context* c, int numffts, int reps, double* report, TuneEntry* out);

extern "C" int fx_autotune(fx_context* c, int numffts, int reps, const char* cache_path, double* report)
{
    if (!c) return fail(FX_EINVAL, "null argument");
    if (reps < 1) return fail(FX_EINVAL, "reps=%d", reps);
    // options, timing and tap state are restored on every exit path, and the
    // trial products never stay in the accumulators
    const int saved_timing = c->timing, saved_tap = c->tap;
    const int saved_plan = c->fplan, saved_xsb = c->xsb, saved_xbatch = c->xbatch;
    c->timing = 0;
    c->tap = 0;
    TuneEntry e;
    int rc = autotune_run(c, numffts, reps, report, &e);
    c->timing = saved_timing;
    c->tap = saved_tap;
    if (rc) {
        c->fplan = saved_plan;
        c->xsb = saved_xsb;
        c->xbatch = saved_xbatch;
        const std::string msg = g_err;
        (void)hipStreamSynchronize(c->stream);
        (void)launch_zero(c->d_vacc, c->vacc_bytes, c->stream);
        (void)launch_zero(c->d_counts, sizeof(int) * c->p.nant * c->p.nant, c->stream);
        (void)hipStreamSynchronize(c->stream);
        return fail(rc, "%s", msg.c_str());
    }
    // the trials accumulated into the visibilities: start clean
    HIPCHK(launch_zero(c->d_vacc, c->vacc_bytes, c->stream));
    HIPCHK(launch_zero(c->d_counts, sizeof(int) * c->p.nant * c->p.nant, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const char* path = cache_path ? cache_path : getenv("FX_TUNE_CACHE");
    if (path && *path) return tune_write(path, e);
    return FX_OK;
}

// The autotuner's trials; leaves the chosen plan, tile and launch count set
// and the result in *out.  The caller restores state on failure.
static int autotune_run(fx_context* c, int numffts, int reps, double* report, TuneEntry* out)
{
    std::vector<int> it(numffts > 0 ? numffts : 1);
    for (int k = 0; k < numffts; k++) it[k] = k;
    FStageCall f;
    int rc = fstage_check(c, numffts, it.data(), 1, 0, &f);
    if (rc) return rc;
    rc = iter_upload(c, it.data(), f, c->stream);
    if (rc) return rc;

    // F-stage radix plans and X-engine tiles: candidates timed round-robin,
    // the best of the rounds kept (differences of a few percent must not
    // drown in the spread of single timings)
    constexpr int kRounds = 3;
    const int nplan = fstage_plans(c->p.lg, c->p.iscomplex);
    std::vector<double> fms(nplan, 1e30);
    for (int round = 0; rc == FX_OK && round < kRounds; round++)
        for (int v = 0; rc == FX_OK && v < nplan; v++) {
            c->fplan = v;
            double ms = 0.0;
            rc = tune_time(c, reps, &ms, [&] { return fstage_enqueue(c, f, c->stream); });
            if (rc == FX_OK) fms[v] = std::min(fms[v], ms);
        }
    int best_plan = 0;
    for (int v = 1; v < nplan; v++)
        if (fms[v] < fms[best_plan]) best_plan = v;
    double best_f = fms[best_plan];
    c->fplan = best_plan;
    // X-engine tiles over up to 8 subints of slots (the K depth of a batched
    // launch: a tile's K pipeline is not in its steady state over one subint)
    int kt = 1;
    while (kt < 8 && (long long)2 * kt * f.npad <= c->max_ffts) kt *= 2;
    const int tiles[3] = {1, 2, 4};
    double xms[3] = {1e30, 1e30, 1e30};
    for (int round = 0; rc == FX_OK && round < kRounds; round++)
        for (int t = 0; rc == FX_OK && t < 3; t++) {
            if (tiles[t] > 1 && c->p.nblk < tiles[t] - 1) continue;   // tile far larger than the signal count
            c->xsb = tiles[t];
            double ms = 0.0;
            rc = tune_time(c, reps, &ms, [&] { return xstage_enqueue(c, 0, kt * f.npad, c->stream); });
            if (rc == FX_OK) xms[t] = std::min(xms[t], ms / kt);
        }
    int bt = 0;
    for (int t = 1; t < 3; t++)
        if (xms[t] < xms[bt]) bt = t;
    // tile and subints per X-engine launch together, in whole steps of k
    // F-stages and one X-engine launch over their slots: a tile also changes
    // the F-stages' speed (what the X-engine leaves in the caches, the power
    // it draws), and a deeper tile needs more subints per launch to reach its
    // steady state, neither of which timing the X-engine alone shows.  Tiles
    // within 25% of the best alone, k = 1, 2, 4, ... (up to the FFT capacity),
    // round-robin over the rounds with each pair's best kept; then the
    // smallest k whose best tile is within 0.2% of the best time per subint (a
    // larger k only adds buffer memory and latency; a 1% margin kept 4 subints
    // per launch at 16 antennas x 2048 channels where 16 ran 7% faster) and
    // that tile.
    int counts[8], nk = 0;
    for (int k = 1; nk < 8 && (long long)k * f.npad <= c->max_ffts; k *= 2) counts[nk++] = k;
    // k subints as callers with a buffer of several subints run them: one
    // delay / F-stage launch over their FFTs (here the trial's FFT list k
    // times, one upload of the longest list serving every k) and one X-engine
    // launch over the slots; where the input capacity does not hold k
    // subints, one F-stage launch per subint
    const FxPlan& pp = c->p;
    auto batch_fits = [&](int k) {
        const long long n = (long long)k * numffts;
        const long long need = pp.iscomplex ? 4LL * n * pp.F : (n * pp.F + 1) / 2;
        return k > 1 && n <= c->max_ffts && need <= c->cap_bytes && (n * pp.F) <= 0x7fffffffLL - 2LL * pp.F;
    };
    FStageCall fk[8];
    int kmax = 0;
    for (int i = 0; i < nk; i++)
        if (batch_fits(counts[i])) kmax = counts[i];
    if (rc == FX_OK && kmax > 1) {
        std::vector<int> itk((size_t)kmax * numffts);
        for (size_t q = 0; q < itk.size(); q++) itk[q] = (int)(q % numffts);
        for (int i = 0; rc == FX_OK && i < nk; i++) {
            if (!batch_fits(counts[i])) continue;
            rc = fstage_check(c, counts[i] * numffts, itk.data(), 1, 0, &fk[i]);
        }
        FStageCall fmax;
        if (rc == FX_OK) rc = fstage_check(c, kmax * numffts, itk.data(), 1, 0, &fmax);
        if (rc == FX_OK) rc = iter_upload(c, itk.data(), fmax, c->stream);
        if (rc == FX_OK) rc = hipStreamSynchronize(c->stream) == hipSuccess ? FX_OK : fail(FX_EHIP, "stream sync");
        for (int i = 0; i < nk; i++) {
            fk[i].contig = 0;                             // the uploaded list's prefix
            fk[i].it0 = 0;
        }
    }
    auto step_ms = [&](int i, double* ms) {
        const int k = counts[i];
        return tune_time(c, reps, ms, [&] {
            if (batch_fits(k)) {
                const int r = fstage_enqueue(c, fk[i], c->stream);
                if (r) return r;
                return xstage_enqueue(c, 0, fk[i].npad, c->stream);
            }
            for (int s = 0; s < k; s++) {
                FStageCall fs = f;
                fs.fft0 = s * f.npad;
                const int r = fstage_enqueue(c, fs, c->stream);
                if (r) return r;
            }
            return xstage_enqueue(c, 0, k * f.npad, c->stream);
        });
    };
    double per_sub[3][8];
    for (int t = 0; t < 3; t++)
        for (int i = 0; i < 8; i++) per_sub[t][i] = 1e30;
    for (int round = 0; rc == FX_OK && round < kRounds; round++)
        for (int t = 0; rc == FX_OK && t < 3; t++) {
            if (xms[t] > 1.25 * xms[bt]) continue;              // far slower alone (or skipped): not a candidate
            c->xsb = tiles[t];
            for (int i = 0; rc == FX_OK && i < nk; i++) {
                double ms = 0.0;
                rc = step_ms(i, &ms);
                if (rc == FX_OK) per_sub[t][i] = std::min(per_sub[t][i], ms / counts[i]);
            }
        }
    int best_k = 1, best_tile = tiles[bt];
    if (rc == FX_OK && nk > 0) {
        double m = 1e30;
        for (int t = 0; t < 3; t++)
            for (int i = 0; i < nk; i++) m = std::min(m, per_sub[t][i]);
        bool found = false;
        for (int i = 0; i < nk && !found; i++) {
            int tb = -1;
            for (int t = 0; t < 3; t++)
                if (per_sub[t][i] <= 1.002 * m && (tb < 0 || per_sub[t][i] < per_sub[tb][i])) tb = t;
            if (tb >= 0) {
                best_k = counts[i];
                best_tile = tiles[tb];
                found = true;
            }
        }
    }
    if (rc) return rc;
    c->xbatch = best_k;
    c->xsb = best_tile;
    const double best_x = xms[best_tile == 1 ? 0 : best_tile == 2 ? 1 : 2];
    if (report) {
        report[0] = best_plan;
        report[1] = best_tile;
        report[2] = best_k;
        report[3] = best_f;
        report[4] = best_x;
    }
    *out = TuneEntry{device_arch(), c->p.nant, c->p.nchan, c->prec, best_plan, best_tile, best_k, best_f, best_x};
    return FX_OK;
}

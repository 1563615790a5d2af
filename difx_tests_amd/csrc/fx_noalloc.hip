// fx_noalloc.hip -- fx_process_noalloc: one subint on caller-owned device
// buffers, every intermediate in one fixed device array, no allocation.
//
// The same kernels as fx_process (delay / rotator tables, fused F-stage,
// X-engine, valid-FFT counts, export), fed from `scratch` below instead of a
// context's buffers.  The FFTs of the call go through in chunks sized to what
// the array holds after the visibility accumulators; each chunk's X-engine
// launch adds into the accumulators, so results equal fx_process to FP32
// rounding (the partial sums are added in a different grouping).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "difxfx.h"
#include "fx_common.h"
#include "fx_kernels.h"

static constexpr size_t kScratchBytes = (size_t)FX_NOALLOC_SCRATCH_MB << 20;

// The scratch array, one copy per GPU, held until the process ends.  Host
// code never uses this name as an address: it is a host-side stand-in; the
// device address comes from hipGetSymbolAddress once per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];      // device address of `scratch` per device

// Device address of this device's scratch array (set up on the first call on
// each device, with that device current).
int scratch_ptr(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fx::set_error(FX_EHIP, hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fx::set_error(FX_EUNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[256];
            snprintf(msg, sizeof(msg), "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return fx::set_error(FX_EHIP, msg);
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return FX_OK;
}

// Twiddles exp(-2 pi i m / F) and the fractional-delay channel frequencies
// (kernel.c:164-165), FP64 arithmetic rounded to f32 as fx_create does.
__global__ void k_tables(int F, int stride, int nfstep, int nchan, double bw, float2* tw, float* subchf,
                         float* stepchf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double PI = 3.14159265358979323846;
    if (t < F) {
        const double ang = -2.0 * M_PI * (double)t / (double)F;
        tw[t] = make_float2((float)cos(ang), (float)sin(ang));
    }
    if (t < stride) subchf[t] = (float)((2 * PI * t * bw) / (double)nchan);
    if (t < nfstep) stepchf[t] = (float)((2 * PI * t * stride * bw) / (double)nchan);
}

__global__ void k_fill(float* a, long long n, float v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// counts[j*nant + k] -> cross_counts[b] (j < k, j-major) and auto_counts[j]
__global__ void k_counts_out(int nant, const int* counts, int* cross_counts, int* auto_counts)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nbl = nant * (nant - 1) / 2;
    if (t < nbl) {
        int b = t, j = 0;
        while (b >= nant - 1 - j) {
            b -= nant - 1 - j;
            j++;
        }
        if (cross_counts) cross_counts[t] = counts[j * nant + j + 1 + b];
    } else if (t < nbl + nant && auto_counts) {
        const int j = t - nbl;
        auto_counts[j] = counts[j * nant + j];
    }
}

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

#define CHK(expr)                                                   \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess) {                                     \
            char m_[256];                                           \
            snprintf(m_, sizeof(m_), "%s: %s", #expr, hipGetErrorString(e_)); \
            return fx::set_error(FX_EHIP, m_);                      \
        }                                                           \
    } while (0)

}  // namespace

extern "C" int fx_process_noalloc(const fx_params* prm, const uint8_t* data, long long stride, const double* delays,
                                  const double* fileoffsets, const int* iteration, int numffts, int split,
                                  float* cross, int* cross_counts, float* autos, int* auto_counts, void* stream)
{
    using namespace fx;
    if (!prm || !data || !delays || !fileoffsets || !iteration) return set_error(FX_EINVAL, "null argument");
    FxPlan p;
    int rc = plan_from_params(prm, &p);
    if (rc) return rc;
    if (prm->prec < FX_PREC_F32 || prm->prec > FX_PREC_F32V) return set_error(FX_EINVAL, "bad precision");
    if (numffts < 1 || split < 1) return set_error(FX_EINVAL, "numffts and split must be >= 1");
    const long long maxoffset = ((long long)numffts * split - 1) * p.F;
    if (maxoffset > 0x7fffffffLL - 2LL * p.F) return set_error(FX_EINVAL, "subint too long for 32-bit offsets");
    // the F-stage reads samples [offset, offset + F) with offset <= maxoffset,
    // real input as aligned 32-bit words up to 4 bytes past the last sample's
    const long long need = p.iscomplex ? 4LL * numffts * split * p.F : ((long long)numffts * split * p.F + 1) / 2;
    if (stride < need + 16 || stride % 16 || ((uintptr_t)data & 15))
        return set_error(FX_EINVAL, "data must be 16-byte aligned with a per-antenna stride that is a multiple of "
                                    "16 and at least the subint's bytes + 16");
    const int fprec = prm->prec == FX_PREC_F32V ? FX_PREC_F32 : prm->prec;
    const int G = fx_fstage_group_size(fprec), kc = xengine_chunk_groups(prm->prec), gran = G * kc;
    p.G = G;
    const int npad = (numffts + gran - 1) / gran * gran;

    // fixed part of the scratch array
    const size_t vacc_bytes = (size_t)p.nchan * p.nbp * FX_BLK * FX_BLK * sizeof(float2);
    const size_t fixed = up256(sizeof(float2) * p.F) + up256(sizeof(float) * p.stride) +
                         up256(sizeof(float) * p.nfstep) + up256(vacc_bytes) +
                         up256(sizeof(int) * (size_t)p.nant * p.nant) +
                         (fprec == FX_PREC_F8 ? 2 * up256(sizeof(float) * (size_t)p.nchan * p.Spad) : 0);
    // per chunk of K FFTs: delay entries, validity bytes, rotator tables, X
    // input, and at F = 32768 the F-stage's row staging
    auto per_chunk = [&](long long K) {
        return up256(sizeof(FxDelay) * K * p.nant) + up256((size_t)K * p.nant) +
               up256(sizeof(float2) * K * p.nant * fx_rtab_len(p)) + up256((size_t)(K / G / kc) * p.nchan * fx_xin_pitch(kc, p.Spad) * 16) +
               up256(fx_fstage_stage_bytes(p.lg, fprec, p.nant, (int)(K / G)));
    };
    long long K = npad;
    while (K > gran && fixed + per_chunk(K) > kScratchBytes) K -= gran;
    if (fixed + per_chunk(K) > kScratchBytes) {
        char msg[256];
        snprintf(msg, sizeof(msg), "workload needs %zu bytes of scratch, the no-allocation array holds %zu",
                 (size_t)(fixed + per_chunk(gran)), kScratchBytes);
        return set_error(FX_ENOMEM, msg);
    }

    void* base = nullptr;
    rc = scratch_ptr(&base);
    if (rc) return rc;
    unsigned char* cur = static_cast<unsigned char*>(base);
    auto take = [&](size_t bytes) {
        unsigned char* r = cur;
        cur += up256(bytes);
        return r;
    };
    float2* tw = reinterpret_cast<float2*>(take(sizeof(float2) * p.F));
    float* subchf = reinterpret_cast<float*>(take(sizeof(float) * p.stride));
    float* stepchf = reinterpret_cast<float*>(take(sizeof(float) * p.nfstep));
    float2* vacc = reinterpret_cast<float2*>(take(vacc_bytes));
    int* counts = reinterpret_cast<int*>(take(sizeof(int) * (size_t)p.nant * p.nant));
    float *scale = nullptr, *scale_inv = nullptr;
    if (fprec == FX_PREC_F8) {
        scale = reinterpret_cast<float*>(take(sizeof(float) * (size_t)p.nchan * p.Spad));
        scale_inv = reinterpret_cast<float*>(take(sizeof(float) * (size_t)p.nchan * p.Spad));
    }
    FxDelay* dtab = reinterpret_cast<FxDelay*>(take(sizeof(FxDelay) * K * p.nant));
    uint8_t* valid = take((size_t)K * p.nant);
    float2* rtab = reinterpret_cast<float2*>(take(sizeof(float2) * K * p.nant * fx_rtab_len(p)));
    const size_t xin_bytes = (size_t)(K / G / kc) * p.nchan * fx_xin_pitch(kc, p.Spad) * 16;
    void* xin = take(xin_bytes);
    void* stage = take(fx_fstage_stage_bytes(p.lg, fprec, p.nant, (int)(K / G)));

    hipStream_t s = static_cast<hipStream_t>(stream);
    // The array keeps the previous call's contents: clear what accumulates and
    // the X input's padding rows (the F-stage writes only real signals' rows)
    CHK(launch_zero(vacc, vacc_bytes, s));
    CHK(launch_zero(counts, sizeof(int) * (size_t)p.nant * p.nant, s));
    CHK(launch_zero(xin, xin_bytes, s));
    {
        const int n = p.F > p.stride ? p.F : p.stride;
        k_tables<<<(n + 255) / 256, 256, 0, s>>>(p.F, p.stride, p.nfstep, p.nchan, p.bw, tw, subchf, stepchf);
        CHK(hipGetLastError());
    }
    if (fprec == FX_PREC_F8) {
        // default FP8 scales of fx_create: one sigma of each component -> 16
        const float sigma = (float)sqrt(p.iscomplex ? p.F * 2.0 * 24.0 * 24.0 / 2.0 : p.F * 4.3128 / 2.0);
        const long long n = (long long)p.nchan * p.Spad;
        k_fill<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(scale, n, sigma / 16.0f);
        k_fill<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(scale_inv, n, 16.0f / sigma);
        CHK(hipGetLastError());
    }
    const int step_q = fstage_step_q(p.lg, 0, p.iscomplex);
    for (int k0 = 0; k0 < numffts; k0 += (int)K) {
        const int nck = numffts - k0 < K ? numffts - k0 : (int)K;
        const int padc = (nck + gran - 1) / gran * gran;
        CHK(launch_delay_rotators(p, delays, fileoffsets, iteration + k0, 0, nck, padc, (int)maxoffset, nullptr, 0,
                                  subchf, stepchf, step_q, dtab, rtab, valid, s));
        FStageArgs a;
        a.p = p;
        a.data = data;
        a.data_stride = stride;
        a.delays = dtab;
        a.rtab = rtab;
        a.tw = tw;
        a.subchf = subchf;
        a.stepchf = stepchf;
        a.out = xin;
        a.kc = kc;
        a.fg0 = 0;
        a.ngroups = padc / G;
        a.scale_inv = scale_inv;
        a.tap = nullptr;
        a.ltap = nullptr;
        a.stage = stage;
        a.plan = 0;
        CHK(launch_fstage(a, fprec, s));
        XArgs x;
        x.p = p;
        x.xin = xin;
        x.fg0 = 0;
        x.ngroups = padc / G;
        x.vacc = vacc;
        x.accumulate = 1;
        x.scale = scale;
        x.sb = 0;
        CHK(launch_xengine(x, prm->prec, s));
        CHK(launch_counts(p.nant, padc, valid, counts, s));
    }
    CHK(launch_export(p, vacc, reinterpret_cast<float2*>(cross), reinterpret_cast<float2*>(autos), nullptr, s));
    if (cross_counts || auto_counts) {
        const int n = p.nant * (p.nant - 1) / 2 + p.nant;
        k_counts_out<<<(n + 255) / 256, 256, 0, s>>>(p.nant, counts, cross_counts, auto_counts);
        CHK(hipGetLastError());
    }
    (void)npad;
    return FX_OK;
}

// fx_single.hip -- the fine-grained kernel.h functions (one antenna, one FFT,
// host arrays in and out) as single-shot GPU calls.
//
// These mirror kernel.c:373-568 and kernel.c:829-854 on caller-owned host
// buffers; they are not the hot path (processAntennasAndBaselinePara is) and
// favour simplicity: each call copies its operands to the device, runs one
// small kernel and copies the result back.  The FxKernel scratch argument of
// the reference is not needed (the rotator tables live on the device) and is
// left untouched.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fx_common.h"
#include "fx_kernels.h"
#include "fx_single.h"
#include "kernel.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

static int g_fftchannels = 0;
static double g_bw = 0.0;

[[noreturn]] void die_hip(hipError_t e, const char* what)
{
    fprintf(stderr, "Error: %s: %s\n", what, hipGetErrorString(e));
    exit(1);
}

#define CK(expr)                                  \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess) die_hip(e_, #expr); \
    } while (0)

// Device scratch freed at scope exit.
struct DBuf {
    void* p = nullptr;
    explicit DBuf(size_t bytes) { CK(hipMalloc(&p, bytes ? bytes : 16)); }
    ~DBuf() { (void)hipFree(p); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

__device__ __forceinline__ float2 cmul_exact(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 sincos_f32(float arg)
{
    double s, c;
    sincos((double)arg, &s, &c);
    return make_float2((float)c, (float)s);
}

// kernel.c:397-426: u_p[i] = (L[code(offset+i, p)], 0); q0 = offset & 1
__global__ void k_unpack(const uint8_t* bytes, int q0, int nsamp, float2* u0, float2* u1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsamp) return;
    const float hi = 3.3359f;
    const float lv[4] = {-hi, -1.0f, 1.0f, hi};
    const int q = q0 + i;
    const int b = bytes[q >> 1], t = q & 1;
    u0[i] = make_float2(lv[(b >> (4 * t)) & 3], 0.0f);
    u1[i] = make_float2(lv[(b >> (4 * t + 2)) & 3], 0.0f);
}

// kernel.c:428-497: u_p[n] = rot[n] * u_p[n], rot[k*sub+s] = rsub[s]*rstep[k]
__global__ void k_fringe(float2* u0, float2* u1, double a, double b, int sub, int stride, int F, double lo,
                         int fracLO)
{
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= F) return;
    const int integerdelay = (int)b;
    const double bf = b - integerdelay;
    const int s = n % sub, k = n / sub;
    double subxval = ((double)s / (double)F) * a;
    subxval += bf;
    double ph = subxval * lo;
    if (fracLO) ph += (lo - (double)(int)lo) * (double)integerdelay;
    const float2 rs = sincos_f32((float)((-2.0 * kPi) * (ph - (double)(int)ph)));
    float2 rot = rs;
    if (k > 0) {
        const double sph = ((((double)k * (double)stride) / (double)F) * a) * lo;
        rot = cmul_exact(rs, sincos_f32((float)((-2.0 * kPi) * (sph - (double)(int)sph))));
    }
    u0[n] = cmul_exact(rot, u0[n]);
    u1[n] = cmul_exact(rot, u1[n]);
}

// kernel.c:512-555 on channels < nchan
__global__ void k_frac(float2* y0, float2* y1, float fd32, int stride, int nchan, double bw)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchan) return;
    const int s = c % stride, k = c / stride;
    const float subchf = (float)((2.0 * kPi * s * bw) / (double)nchan);
    float2 fr = sincos_f32(subchf * fd32);
    if (k > 0) {
        const float stepchf = (float)((2.0 * kPi * k * stride * bw) / (double)nchan);
        fr = cmul_exact(fr, sincos_f32(stepchf * fd32));
    }
    y0[c] = cmul_exact(fr, y0[c]);
    y1[c] = cmul_exact(fr, y1[c]);
}

__global__ void k_conj(const float2* y, float2* z, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = make_float2(y[i].x, -y[i].y);
}

// kernel.c:829-854 for one FFT: chan/conj are [nant][2][nchan], pair list
// holds (j, k, b) of the valid baselines.
__global__ void k_xacc(const float2* chan, const float2* conjc, const int4* pairs, int npairs, int nchan,
                       float2* vis)
{
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npairs * 4 * nchan) return;
    const int c = (int)(t % nchan);
    const int prod = (int)((t / nchan) & 3);
    const int4 pr = pairs[t / nchan / 4];
    const int l = prod >> 1, m = prod & 1;
    const float2 x = chan[((long long)pr.x * 2 + l) * nchan + c];
    const float2 y = conjc[((long long)pr.y * 2 + m) * nchan + c];
    const float2 p = make_float2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
    float2* acc = vis + ((long long)pr.w * 4 + prod) * nchan + c;
    *acc = make_float2(acc->x + p.x, acc->y + p.y);
}

// kernel.c:373-386
__global__ void k_station_delay(const double* c, int fftindex, double* out)
{
#pragma clang fp contract(off)
    const double d0 = c[0] * fftindex * fftindex + c[1] * fftindex + c[2];
    const double d1 = c[0] * (fftindex + 0.5) * (fftindex + 0.5) + c[1] * (fftindex + 0.5) + c[2];
    const double d2 = c[0] * (fftindex + 1.0) * (fftindex + 1.0) + c[1] * (fftindex + 1.0) + c[2];
    const double a = d2 - d0;
    const double b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
    out[0] = a * 0.5 + b;
    out[1] = a;
    out[2] = b;
}

void h2d(void* d, const void* h, size_t n) { CK(hipMemcpy(d, h, n, hipMemcpyHostToDevice)); }
void d2h(void* h, const void* d, size_t n) { CK(hipMemcpy(h, d, n, hipMemcpyDeviceToHost)); }

}  // namespace

// Per-antenna buffers of the reference's FFT loop (kernel.c:725-800) after
// processAntennasAndBaselinePara: the loop overwrites unpacked[j],
// channelised[j] and conjchannels[j] at every FFT in which antenna j is
// valid, so the state it leaves is that of each antenna's last valid FFT.
// Only those FFTs are formed here, on the device from the call's packed input
// (d_data, antenna a at d_data + a*stride): unpacked and fringe-rotated
// samples (F), the spectrum (F bins, the lower nchan frac-corrected) and the
// conjugate of the lower nchan.  Antennas with no valid FFT keep their
// buffers, as the reference's do.
void fxk_fill_antenna_buffers(const uint8_t* d_data, long long stride, const FxDelay* last, const int* has,
                              int nant, int F, int nchan, int sub, int stridesize, double lo, int fracLO, double bw,
                              cf32*** unpacked, cf32*** channelised, cf32*** conjchannels)
{
    std::vector<int> ants;
    for (int j = 0; j < nant; j++)
        if (has[j]) ants.push_back(j);
    if (ants.empty()) return;
    const int nv = (int)ants.size();
    const int lg = fx_ilog2_exact(F);
    if (lg < 0 || !fx::fstage_supported(lg)) {
        fprintf(stderr, "Error calling FFT");
        exit(1);
    }
    std::vector<float2> tw(F);
    for (int m = 0; m < F; m++) {
        const double ang = -2.0 * kPi * (double)m / (double)F;
        tw[m] = make_float2((float)cos(ang), (float)sin(ang));
    }
    DBuf u(sizeof(float2) * 2 * (size_t)F * nv), y(sizeof(float2) * 2 * (size_t)F * nv),
        z(sizeof(float2) * 2 * (size_t)nchan * nv), dtw(sizeof(float2) * F);
    h2d(dtw.p, tw.data(), sizeof(float2) * F);
    for (int v = 0; v < nv; v++) {
        const FxDelay& e = last[ants[v]];
        float2* u0 = u.as<float2>() + (size_t)(2 * v) * F;
        k_unpack<<<(F + 255) / 256, 256>>>(d_data + (size_t)ants[v] * stride + e.offset / 2, e.offset & 1, F, u0,
                                           u0 + F);
        k_fringe<<<(F + 255) / 256, 256>>>(u0, u0 + F, e.a, e.b, sub, stridesize, F, lo, fracLO);
    }
    CK(hipGetLastError());
    CK(fx::launch_fft_c2c(u.as<float2>(), y.as<float2>(), dtw.as<float2>(), lg, 2 * nv, 0));
    for (int v = 0; v < nv; v++) {
        float2* y0 = y.as<float2>() + (size_t)(2 * v) * F;
        float2* z0 = z.as<float2>() + (size_t)(2 * v) * nchan;
        k_frac<<<(nchan + 255) / 256, 256>>>(y0, y0 + F, (float)last[ants[v]].fd, stridesize, nchan, bw);
        k_conj<<<(nchan + 255) / 256, 256>>>(y0, z0, nchan);
        k_conj<<<(nchan + 255) / 256, 256>>>(y0 + F, z0 + nchan, nchan);
    }
    CK(hipGetLastError());
    for (int v = 0; v < nv; v++) {
        const int j = ants[v];
        for (int p = 0; p < 2; p++) {
            d2h(unpacked[j][p], u.as<float2>() + (size_t)(2 * v + p) * F, sizeof(float2) * F);
            d2h(channelised[j][p], y.as<float2>() + (size_t)(2 * v + p) * F, sizeof(float2) * F);
            d2h(conjchannels[j][p], z.as<float2>() + (size_t)(2 * v + p) * nchan, sizeof(float2) * nchan);
        }
    }
}

int fxk_fftchannels(void) { return g_fftchannels; }
double fxk_bandwidth(void) { return g_bw; }
void fxk_set_geometry(int fftchannels, double bw)
{
    g_fftchannels = fftchannels;
    g_bw = bw;
}

extern "C" {

void unpackReal2bit(u8* inputdata, cf32** unpacked, int offset, int nsamp)
{
    if (nsamp <= 0) return;
    const int b0 = offset / 2, b1 = (offset + nsamp - 1) / 2;
    DBuf bytes((size_t)(b1 - b0 + 1)), u0(sizeof(float2) * nsamp), u1(sizeof(float2) * nsamp);
    h2d(bytes.p, inputdata + b0, (size_t)(b1 - b0 + 1));
    k_unpack<<<(nsamp + 255) / 256, 256>>>(bytes.as<uint8_t>(), offset & 1, nsamp, u0.as<float2>(),
                                           u1.as<float2>());
    CK(hipGetLastError());
    d2h(unpacked[0], u0.p, sizeof(float2) * nsamp);
    d2h(unpacked[1], u1.p, sizeof(float2) * nsamp);
}

void unpack(u8* inputdata, cf32** unpacked, int offset, int nbit, int fftchannels)
{
    // kernel.c:388-395
    if (nbit == 2) {
        unpackReal2bit(inputdata, unpacked, offset, fftchannels);
    } else {
        fprintf(stderr, "Unsupported number of bits!!!");
    }
}

void getStationDelay(int antenna, int fftindex, double* meandelay, double* a, double* b, double** delays)
{
    DBuf c(3 * sizeof(double)), o(3 * sizeof(double));
    h2d(c.p, delays[antenna], 3 * sizeof(double));
    k_station_delay<<<1, 1>>>(c.as<double>(), fftindex, o.as<double>());
    CK(hipGetLastError());
    double r[3];
    d2h(r, o.p, sizeof(r));
    *meandelay = r[0];
    *a = r[1];
    *b = r[2];
}

void fringerotate(FxKernel* kernel, cf32** unpacked, f64 a, f64 b, int substridesize, int stridesize, double lo,
                  int fftchannels, double fractionalLoFreq)
{
    (void)kernel;
    const int F = fftchannels;
    if (substridesize <= 0 || F % substridesize) {
        fprintf(stderr, "Error: fftchannels must be a multiple of substridesize\n");
        exit(1);
    }
    DBuf u0(sizeof(float2) * F), u1(sizeof(float2) * F);
    h2d(u0.p, unpacked[0], sizeof(float2) * F);
    h2d(u1.p, unpacked[1], sizeof(float2) * F);
    k_fringe<<<(F + 255) / 256, 256>>>(u0.as<float2>(), u1.as<float2>(), a, b, substridesize, stridesize, F, lo,
                                       fractionalLoFreq != 0.0);
    CK(hipGetLastError());
    d2h(unpacked[0], u0.p, sizeof(float2) * F);
    d2h(unpacked[1], u1.p, sizeof(float2) * F);
}

void dofft(FxKernel* kernel, cf32** unpacked, cf32** channelised)
{
    (void)kernel;
    const int F = fxk_fftchannels();
    const int lg = fx_ilog2_exact(F);
    if (lg < 0 || !fx::fstage_supported(lg)) {
        fprintf(stderr, "Error calling FFT");
        exit(1);
    }
    std::vector<float2> tw(F);
    for (int m = 0; m < F; m++) {
        const double ang = -2.0 * kPi * (double)m / (double)F;
        tw[m] = make_float2((float)cos(ang), (float)sin(ang));
    }
    DBuf in(sizeof(float2) * 2 * F), out(sizeof(float2) * 2 * F), dtw(sizeof(float2) * F);
    h2d(in.p, unpacked[0], sizeof(float2) * F);
    h2d(in.as<float2>() + F, unpacked[1], sizeof(float2) * F);
    h2d(dtw.p, tw.data(), sizeof(float2) * F);
    CK(fx::launch_fft_c2c(in.as<float2>(), out.as<float2>(), dtw.as<float2>(), lg, 2, 0));
    d2h(channelised[0], out.p, sizeof(float2) * F);
    d2h(channelised[1], out.as<float2>() + F, sizeof(float2) * F);
}

void fracSampleCorrect(FxKernel* kernel, cf32** channelised, f64 fracdelay, int stridesize, int nchan)
{
    (void)kernel;
    DBuf y0(sizeof(float2) * nchan), y1(sizeof(float2) * nchan);
    h2d(y0.p, channelised[0], sizeof(float2) * nchan);
    h2d(y1.p, channelised[1], sizeof(float2) * nchan);
    // kernel.c:164-165: channel frequencies of the bandwidth given to init_FxKernel
    k_frac<<<(nchan + 255) / 256, 256>>>(y0.as<float2>(), y1.as<float2>(), (float)fracdelay, stridesize, nchan,
                                         fxk_bandwidth());
    CK(hipGetLastError());
    d2h(channelised[0], y0.p, sizeof(float2) * nchan);
    d2h(channelised[1], y1.p, sizeof(float2) * nchan);
}

void conjChannels(cf32** channelised, cf32** conjchannels, int nchan)
{
    DBuf y(sizeof(float2) * 2 * nchan), z(sizeof(float2) * 2 * nchan);
    h2d(y.p, channelised[0], sizeof(float2) * nchan);
    h2d(y.as<float2>() + nchan, channelised[1], sizeof(float2) * nchan);
    k_conj<<<(2 * nchan + 255) / 256, 256>>>(y.as<float2>(), z.as<float2>(), 2 * nchan);
    CK(hipGetLastError());
    d2h(conjchannels[0], z.p, sizeof(float2) * nchan);
    d2h(conjchannels[1], z.as<float2>() + nchan, sizeof(float2) * nchan);
}

void processBaselinePara(int* nant, int* antValid, cf32*** channelised, cf32*** conjchannels,
                         cf32*** visibilities, int* nchan, int* baselineCount)
{
    const int na = *nant, nc = *nchan;
    std::vector<int4> pairs;
    int b = 0;
    for (int j = 0; j < na; j++)
        for (int k = j + 1; k < na; k++, b++)
            if (antValid[j] && antValid[k]) pairs.push_back(make_int4(j, k, 0, b));
    if (pairs.empty()) return;
    const int nbl = na * (na - 1) / 2;
    DBuf ch(sizeof(float2) * na * 2 * nc), cj(sizeof(float2) * na * 2 * nc),
        vis(sizeof(float2) * (size_t)nbl * 4 * nc), pr(sizeof(int4) * pairs.size());
    for (int j = 0; j < na; j++)
        for (int p = 0; p < 2; p++) {
            if (!antValid[j]) continue;
            h2d(ch.as<float2>() + ((size_t)j * 2 + p) * nc, channelised[j][p], sizeof(float2) * nc);
            h2d(cj.as<float2>() + ((size_t)j * 2 + p) * nc, conjchannels[j][p], sizeof(float2) * nc);
        }
    for (int i = 0; i < nbl; i++)
        for (int p = 0; p < 4; p++) h2d(vis.as<float2>() + ((size_t)i * 4 + p) * nc, visibilities[i][p], sizeof(float2) * nc);
    h2d(pr.p, pairs.data(), sizeof(int4) * pairs.size());
    const long long n = (long long)pairs.size() * 4 * nc;
    k_xacc<<<(unsigned)((n + 255) / 256), 256>>>(ch.as<float2>(), cj.as<float2>(), pr.as<int4>(),
                                                  (int)pairs.size(), nc, vis.as<float2>());
    CK(hipGetLastError());
    for (const int4& q : pairs) {
        for (int p = 0; p < 4; p++) d2h(visibilities[q.w][p], vis.as<float2>() + ((size_t)q.w * 4 + p) * nc, sizeof(float2) * nc);
        baselineCount[q.w]++;
    }
}

void processBaseline(int* nant, int* antValid, cf32*** channelised, cf32*** conjchannels, cf32*** visibilities,
                     int* nchan, int* baselineCount)
{
    // kernel.c:256-281 is the same computation as processBaselinePara
    processBaselinePara(nant, antValid, channelised, conjchannels, visibilities, nchan, baselineCount);
}

void stationDelayAndOffset(double** delays, int* iteration, double* antFileOffsets, int* antValid, int* offset,
                           double* fracDelay, double* delaya, double* delayb, double* sampletime, int* fftchannels,
                           int numffts)
{
    // kernel.c:938-966 (antenna 0: the signature carries no antenna index)
    FxPlan p;
    memset(&p, 0, sizeof(p));
    p.nant = 1;
    p.F = *fftchannels;
    p.Ts = *sampletime;
    DBuf model(4 * sizeof(double)), it(2 * sizeof(int)), tab(sizeof(FxDelay));
    double m[4] = {delays[0][0], delays[0][1], delays[0][2], antFileOffsets[0]};
    h2d(model.p, m, sizeof(m));
    const int io[2] = {*iteration, 0};                    // FFT index, buffer origin 0
    h2d(it.p, io, sizeof(io));
    CK(fx::launch_delay(p, model.as<double>(), model.as<double>() + 3, it.as<int>(), 1, 1,
                        (numffts - 1) * *fftchannels, it.as<int>() + 1, tab.as<FxDelay>(), 0));
    FxDelay e;
    d2h(&e, tab.p, sizeof(e));
    *delaya = e.a;
    *delayb = e.b;
    *offset = e.offset;
    *fracDelay = e.fd;
    antValid[0] = e.valid;
}

void processAntennas(FxKernel* kernel, int* nant, int* numffts, int* fftchannels, double* antFileOffsets,
                     double* sampletime, int* antValid, u8** inputData, cf32*** unpacked, cf32*** channelised,
                     cf32*** conjchannels, double** delays, int* nbit, int* substridesize, int* stridesize,
                     int* fractionalLoFreq, double* lo, int* nchan)
{
    // kernel.c:205-254.  Every FFT overwrites the per-antenna buffers, so the
    // state left behind is that of the last FFT; only it is computed.  The
    // reference's early return on an invalid antenna is not reproduced.
    const int F = *fftchannels, i = *numffts - 1;
    if (i < 0) return;
    FxPlan p;
    memset(&p, 0, sizeof(p));
    p.nant = *nant;
    p.F = F;
    p.Ts = *sampletime;
    DBuf model(4 * sizeof(double) * *nant), it(2 * sizeof(int)), tab(sizeof(FxDelay) * *nant);
    std::vector<double> m(4 * (size_t)*nant);
    for (int j = 0; j < *nant; j++) {
        for (int q = 0; q < 3; q++) m[3 * (size_t)j + q] = delays[j][q];
        m[3 * (size_t)*nant + j] = antFileOffsets[j];
    }
    h2d(model.p, m.data(), sizeof(double) * m.size());
    const int io[2] = {i, 0};                             // FFT index, buffer origin 0
    h2d(it.p, io, sizeof(io));
    CK(fx::launch_delay(p, model.as<double>(), model.as<double>() + 3 * *nant, it.as<int>(), 1, 1,
                        (*numffts - 1) * F, it.as<int>() + 1, tab.as<FxDelay>(), 0));
    std::vector<FxDelay> e((size_t)*nant);
    d2h(e.data(), tab.p, sizeof(FxDelay) * *nant);
    for (int j = 0; j < *nant; j++) {
        antValid[j] = e[j].valid;
        if (!e[j].valid) continue;
        unpack(inputData[j], unpacked[j], e[j].offset, *nbit, F);
        fringerotate(kernel, unpacked[j], e[j].a, e[j].b, *substridesize, *stridesize, *lo, F, *fractionalLoFreq);
        dofft(kernel, unpacked[j], channelised[j]);
        fracSampleCorrect(kernel, channelised[j], e[j].fd, *stridesize, *nchan);
        conjChannels(channelised[j], conjchannels[j], *nchan);
    }
}

void unpackImpl(cf32** unpacked, int split, int* nant, int* fftchannels, u8* inputData, int* offset, int* nbit)
{
    // kernel.c:881-895 without the local reallocation of `unpacked`
    (void)split;
    (void)nant;
    unpack(inputData, unpacked, *offset, *nbit, *fftchannels);
}

void fringeRotateImpl(FxKernel* kernel, cf32** unpacked, f64 a, f64 b, int substridesize, int stridesize, double lo,
                      int fftchannels, double fractionalLoFreq, cf32** unpacked_out, FxKernel* kernel_out)
{
    fringerotate(kernel, unpacked, a, b, substridesize, stridesize, lo, fftchannels, fractionalLoFreq);
    memcpy(unpacked_out, unpacked, 2 * sizeof(cf32*));
    if (kernel && kernel_out) memcpy(kernel_out, kernel, sizeof(FxKernel));
}

void doFFTImpl(FxKernel* kernel, cf32** unpacked, cf32** channelised, int* fftchannels)
{
    // kernel.c:906-915 without reallocating `channelised`
    (void)fftchannels;
    dofft(kernel, unpacked, channelised);
}

void fracSampleCorrectImpl(FxKernel* kernel, cf32** channelised, double* fractionaldelay, int* stridesize, int* nchan,
                           cf32** channelised_out)
{
    fracSampleCorrect(kernel, channelised, *fractionaldelay, *stridesize, *nchan);
    memcpy(channelised_out, channelised, 2 * sizeof(cf32*));
}

void conjChannelsImpl(cf32** channelised, cf32** conjchannels, int* nchan, int* fftchannels, cf32** channelised_out)
{
    // kernel.c:924-936 without reallocating `channelised`
    (void)fftchannels;
    conjChannels(channelised, conjchannels, *nchan);
    memcpy(channelised_out, channelised, 2 * sizeof(cf32*));
}

}  // extern "C"

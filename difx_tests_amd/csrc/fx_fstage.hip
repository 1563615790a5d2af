// fx_fstage.hip -- F-stage of the MI355X FX engine (gfx950).
//
//   fx_delay_kernel   FP64 delay model per (FFT, antenna): kernel.c:373-386
//                     (getStationDelay) and kernel.c:736-756 (integer/
//                     fractional split, edge clamps, validity).
//   fx_delay_rot_kernel  the same delay entry fused with the FP64-phase
//                     fringe and fractional-delay rotator tables per
//                     (FFT, antenna) (kernel.c:428-497, 512-555).
//   fx_fstage_kernel  fused 2-bit unpack (LDS byte window recoded to one
//                     polarisation's code pairs + 32-word LDS level table),
//                     fringe rotation (kernel.c:428-497), forward C2C FFT
//                     (kernel.c:499-510) as LDS radix-8/16 passes with
//                     in-register butterflies (Stockham, in-place
//                     decimation in frequency, or a last radix-2 pass across
//                     half-waves), fractional-sample correction of channels
//                     < nchan (kernel.c:512-555) and the conversion to the
//                     X-engine's corner-turned input (FP32 / FP16 / FP8).
//   fx_fstage_eo_kernel  the same for F = 32768 from two 16384-point halves.
//
// One workgroup handles G consecutive FFTs of one antenna and one
// polarisation, and keeps the G spectra of every channel it owns in
// registers, so each (channel, signal) run of G FFTs leaves in one 16-byte
// store.
#include <atomic>
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "fx_common.h"
#include "fx_kernels.h"

namespace fx {

static constexpr double kPi = 3.14159265358979323846;

// --------------------------------------------------------------------------
// FP64 delay table

// Delay-table entry of antenna j at global FFT index fftindex, for a buffer
// whose first FFT is `origin` (kernel.c:373-386 and 736-756).
__device__ FxDelay delay_entry(const FxPlan& p, const double* __restrict__ delays,
                               const double* __restrict__ fileoff, int j, int fftindex, int origin, int maxoffset)
{
#pragma clang fp contract(off)
    const double c0 = delays[3 * j], c1 = delays[3 * j + 1], c2 = delays[3 * j + 2];
    // kernel.c:378-385
    const double d0 = c0 * fftindex * fftindex + c1 * fftindex + c2;
    const double d1 = c0 * (fftindex + 0.5) * (fftindex + 0.5) + c1 * (fftindex + 0.5) + c2;
    const double d2 = c0 * (fftindex + 1.0) * (fftindex + 1.0) + c1 * (fftindex + 1.0) + c2;
    const double a = d2 - d0;
    const double b = d0 + (d1 - (a * 0.5 + d0)) / 3.0;
    const double mean = a * 0.5 + b;
    // kernel.c:736-756; the offset is relative to a buffer whose first FFT is
    // `origin` (0 for the reference's single-subint buffer)
    const double x = (mean - fileoff[j]) / p.Ts;
    const int n = (int)(x + 0.5);
    double fd = -(x - n) * p.Ts;
    int o = (fftindex - origin) * p.F - n;
    if (o == -1) {
        ++o;
        fd += p.Ts;
    }
    if (o == maxoffset + 1) {
        --o;
        fd -= p.Ts;
    }
    FxDelay e;
    e.a = a;
    e.b = b;
    e.fd = fd;
    e.offset = o;
    e.valid = !(o < 0 || o > maxoffset);
    return e;
}

__global__ void fx_delay_kernel(FxPlan p, const double* __restrict__ delays,
                                const double* __restrict__ fileoff, const int* __restrict__ iteration,
                                int numffts, int nffts_pad, int maxoffset, const int* __restrict__ origin_dev,
                                FxDelay* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nffts_pad * p.nant) return;
    const int k = t / p.nant;
    const int j = t - k * p.nant;
    FxDelay e;
    if (k >= numffts) {
        e.a = 0.0; e.b = 0.0; e.fd = 0.0; e.offset = 0; e.valid = 0;
    } else {
        e = delay_entry(p, delays, fileoff, j, iteration[k], *origin_dev, maxoffset);
    }
    out[t] = e;
}

hipError_t launch_delay(const FxPlan& p, const double* delays, const double* fileoff, const int* iteration,
                        int numffts, int nffts_pad, int maxoffset, const int* origin_dev, FxDelay* out,
                        hipStream_t s)
{
    const int n = nffts_pad * p.nant;
    if (n <= 0) return hipSuccess;
    fx_delay_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, delays, fileoff, iteration, numffts, nffts_pad,
                                                    maxoffset, origin_dev, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// complex helpers

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// a*b with every product and sum rounded separately (the reference's IPP
// complex multiply); used where the oracle must agree bit for bit.
__device__ __forceinline__ float2 cmul_exact(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos(2 pi q / 32) from the first quadrant by symmetry.
__host__ __device__ constexpr float kc32(int q)
{
    constexpr float c[9] = {1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f,
                            7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f,
                            0.0f};
    return q <= 8 ? c[q] : q <= 16 ? -c[16 - q] : q <= 24 ? -c[q - 16] : c[32 - q];
}
__host__ __device__ constexpr float ks32(int q) { return kc32((q + 24) & 31); }  // sin x = cos(x - pi/2)

// x * exp(-2 pi i q / 32); q is a constant once the caller is unrolled.
__device__ __forceinline__ float2 twc(float2 x, int q)
{
    if (q == 0) return x;
    if (q == 8) return make_float2(x.y, -x.x);
    if (q == 16) return make_float2(-x.x, -x.y);
    if (q == 24) return make_float2(-x.y, x.x);
    if (q == 4) {
        const float r = 7.071067812e-01f;
        return make_float2((x.x + x.y) * r, (x.y - x.x) * r);
    }
    if (q == 12) {
        const float r = 7.071067812e-01f;
        return make_float2((x.y - x.x) * r, -(x.x + x.y) * r);
    }
    const float c = kc32(q), s = ks32(q);
    return make_float2(x.x * c + x.y * s, x.y * c - x.x * s);
}

// sin and cos of an FP32 argument, evaluated in FP64 and rounded to FP32 (the
// oracle rounds the C library's double sin/cos the same way).  The rotator
// arguments of a valid antenna are bounded (|x| < 8 rad): the reduction by
// pi/2 is Cody-Waite with a three-part constant (exact for |x| < 2^20), the
// kernels are the fdlibm minimax polynomials on [-pi/4, pi/4].
__device__ __forceinline__ float2 sincos_rounded(float xf)
{
    const double x = (double)xf;
    const double q = rint(x * 6.36619772367581382433e-01);       // 2/pi
    double r = fma(-q, 1.57079632673412561417e+00, x);          // pi/2 high
    r = fma(-q, 6.07710050650619224932e-11, r);                 // pi/2 mid
    r = fma(-q, 2.02226624879595063154e-21, r);                 // pi/2 low
    const double z = r * r;
    const double sp = -1.66666666666666324348e-01 + z * (8.33333333332248946124e-03 +
                      z * (-1.98412698298579493134e-04 + z * (2.75573137070700676789e-06 +
                      z * (-2.50507602534068634195e-08 + z * 1.58969099521155010221e-10))));
    const double cp = 4.16666666666666019037e-02 + z * (-1.38888888888741095749e-03 +
                      z * (2.48015872894767294178e-05 + z * (-2.75573143513906633035e-07 +
                      z * (2.08757232129817482790e-09 + z * -1.13596475577881948265e-11))));
    const double sn = r + r * z * sp;
    const double cs = 1.0 - 0.5 * z + z * z * cp;
    const int quad = ((int)q) & 3;
    double so, co;
    switch (quad) {
    case 0: so = sn; co = cs; break;
    case 1: so = cs; co = -sn; break;
    case 2: so = -sn; co = -cs; break;
    default: so = -cs; co = sn; break;
    }
    return make_float2((float)co, (float)so);
}

// --------------------------------------------------------------------------
// Rotator tables per (FFT, antenna), shared by both polarisations:
//   [0, sub)                      fringe rotator of sample s within a block
//   [sub, sub + nstep)            fringe rotator of block k (k = 0 is exactly 1)
//   [.., + stride)                fractional-delay rotator of channel s
//   [.., + nfstep)                fractional-delay rotator of channel block k
// FP64 phase arguments in the reference's operation order (kernel.c:428-497,
// 512-555; contraction off), sin/cos rounded to FP32.

// Rotator tables of one (FFT, antenna) from its delay entry: sub fine and
// nstep step fringe rotators (FP64 phases, kernel.c:428-497), stride fine and
// nfstep step fractional-delay rotators (kernel.c:512-555); this thread's
// entries first, first + step, ...  F is a power of two, so the reference's
// divisions by F are the exact products with 1/F.
__device__ void rotator_fill(const FxPlan& p, const FxDelay& d, const float* __restrict__ subchf,
                             const float* __restrict__ stepchf, int step_q, float2* __restrict__ out, int first,
                             int step)
{
#pragma clang fp contract(off)
    const int ntab = p.sub + p.nstep + p.stride + p.nfstep;
    // step_q and nstep are powers of two (checked at launch)
    const int qbits = step_q > 0 ? __builtin_ctz(step_q) : 0;
    const int rbits = __builtin_ctz(p.nstep) - qbits;
    const int integerdelay = (int)d.b;
    const double bfrac = d.b - integerdelay;
    const double lofr = p.lofrac * (double)integerdelay;
    const float fd32 = (float)d.fd;
    const double inv_f = 1.0 / (double)p.F;                     // exact
    for (int e = first; e < ntab; e += step) {
        float arg;
        bool unit = false;
        if (e < p.sub) {
            double subxval = ((double)e * inv_f) * d.a;
            subxval += bfrac;
            double ph = subxval * p.lo;
            if (p.fracLO) ph += lofr;
            arg = (float)((-2.0 * kPi) * (ph - (double)(int)ph));
        } else if (e < p.sub + p.nstep) {
            const int k = e - p.sub;
            const double stepxval = (((double)k * (double)p.stride) * inv_f) * d.a;
            const double ph = stepxval * p.lo;
            arg = (float)((-2.0 * kPi) * (ph - (double)(int)ph));
            unit = (k == 0);
        } else if (e < p.sub + p.nstep + p.stride) {
            arg = subchf[e - p.sub - p.nstep] * fd32;
        } else {
            const int k = e - p.sub - p.nstep - p.stride;
            arg = stepchf[k] * fd32;
            unit = (k == 0);
        }
        const float2 v = unit ? make_float2(1.0f, 0.0f) : sincos_rounded(arg);
        out[e] = v;
        if (step_q > 0 && e >= p.sub && e < p.sub + p.nstep) {
            const int k = e - p.sub;
            out[ntab + ((k & (step_q - 1)) << rbits) + (k >> qbits)] = v;
        }
    }
}

// One half-wave per (FFT k of the call, antenna j), 2 * kDelayWaves per
// block: every lane forms the delay entry (global FFT index and buffer
// origin as launch_delay_rotators describes), the half-wave's first lane
// writes it and the FFT slot's
// validity byte, and the half-wave fills the rotator tables.  No barrier: the
// entry is the same in every lane of a half-wave.  Two entries per wave halve
// the FP64 delay-model work per entry, and a table of 160 entries (1024
// channels) takes five full 32-lane rounds instead of three 64-lane rounds
// with the last one half idle.
// Arguments 4 (it0) and 9 (origin) are updated in recorded graphs
// (kDelayArgIt0, kDelayArgOrigin); keep their positions.
static constexpr int kDelayWaves = 4;
__global__ __launch_bounds__(64 * kDelayWaves) void fx_delay_rot_kernel(
    FxPlan p, const double* __restrict__ delays, const double* __restrict__ fileoff, const int* __restrict__ iteration,
    int it0, int numffts, long long nfa, int maxoffset, const int* __restrict__ origin_dev, int origin,
    const float* __restrict__ subchf, const float* __restrict__ stepchf, int step_q, FxDelay* __restrict__ dtab,
    float2* __restrict__ rtab, uint8_t* __restrict__ valid_slots)
{
    const long long fa = (long long)blockIdx.x * (2 * kDelayWaves) + (threadIdx.x >> 5);
    const int lane = threadIdx.x & 31;
    if (fa >= nfa) return;
    const int k = (int)(fa / p.nant), j = (int)(fa - (long long)k * p.nant);
    FxDelay e;
    if (k >= numffts) {
        e.a = 0.0; e.b = 0.0; e.fd = 0.0; e.offset = 0; e.valid = 0;
    } else if (origin_dev) {
        // {origin, first index, list flag} in device memory (graph launches)
        const int g = origin_dev[2] ? iteration[k] : origin_dev[1] + k;
        e = delay_entry(p, delays, fileoff, j, g, origin_dev[0], maxoffset);
    } else {
        const int g = iteration ? iteration[k] : it0 + k;
        e = delay_entry(p, delays, fileoff, j, g, origin, maxoffset);
    }
    if (lane == 0) {
        dtab[fa] = e;
        valid_slots[fa] = (uint8_t)e.valid;
    }
    if (e.valid) rotator_fill(p, e, subchf, stepchf, step_q, rtab + fa * fx_rtab_len(p), lane, 32);
}

const void* delay_rot_kernel_fn() { return reinterpret_cast<const void*>(&fx_delay_rot_kernel); }

hipError_t launch_delay_rotators(const FxPlan& p, const double* delays, const double* fileoff, const int* iteration,
                                 int it0, int numffts, int nffts_pad, int maxoffset, const int* origin_dev, int origin,
                                 const float* subchf, const float* stepchf, int step_q, FxDelay* dtab, float2* rtab,
                                 uint8_t* valid_slots, hipStream_t s)
{
    const long long n = (long long)nffts_pad * p.nant;
    if (n <= 0) return hipSuccess;
    if (step_q < 0 || (step_q & (step_q - 1)) != 0 || (p.nstep & (p.nstep - 1)) != 0 || step_q > p.nstep)
        return hipErrorInvalidValue;
    fx_delay_rot_kernel<<<(unsigned)((n + 2 * kDelayWaves - 1) / (2 * kDelayWaves)), 64 * kDelayWaves, 0, s>>>(
        p, delays, fileoff, iteration, it0, numffts, n, maxoffset, origin_dev, origin, subchf, stepchf, step_q, dtab,
        rtab, valid_slots);
    return hipGetLastError();
}

__host__ __device__ constexpr int cbrev(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

__host__ __device__ constexpr int clog2(int n) { return n <= 1 ? 0 : 1 + clog2(n >> 1); }

// Butterfly (u, v) <- (u + w v, u - w v), w = exp(-2 pi i q / 32); q and
// only_sum are constants once the caller is unrolled.  A non-trivial w is
// applied in factored form, w v = f * (a, b) with f the larger of cos and sin
// in magnitude and a, b one FMA each with the ratio t of the other to it, so
// that each output component is one more FMA: 6 instructions per butterfly
// instead of 8 (the FMA butterfly of Linzer and Feig, 1993).
__device__ __forceinline__ void bfly_tw(float2& u, float2& v, int q, bool only_sum)
{
    float2 p, m;
    if (q == 0) {
        p = cadd(u, v);
        m = csub(u, v);
    } else if (q == 8) {
        p = make_float2(u.x + v.y, u.y - v.x);
        m = make_float2(u.x - v.y, u.y + v.x);
    } else if (q == 16) {
        p = csub(u, v);
        m = cadd(u, v);
    } else if (q == 24) {
        p = make_float2(u.x - v.y, u.y + v.x);
        m = make_float2(u.x + v.y, u.y - v.x);
    } else {
        const float c = kc32(q), s = ks32(q);
        float a, b, f;
        if (__builtin_fabsf(c) >= __builtin_fabsf(s)) {      // w v = c (v.x + t v.y, v.y - t v.x), t = s / c
            const float t = s / c;
            a = __builtin_fmaf(v.y, t, v.x);
            b = __builtin_fmaf(-v.x, t, v.y);
            f = c;
        } else {                                            // w v = s (t v.x + v.y, t v.y - v.x), t = c / s
            const float t = c / s;
            a = __builtin_fmaf(v.x, t, v.y);
            b = __builtin_fmaf(v.y, t, -v.x);
            f = s;
        }
        p = make_float2(__builtin_fmaf(a, f, u.x), __builtin_fmaf(b, f, u.y));
        m = make_float2(__builtin_fmaf(a, -f, u.x), __builtin_fmaf(b, -f, u.y));
    }
    u = p;
    if (!only_sum) v = m;
}

// In-register forward DFT of size R <= 32, natural order in and out (radix-2
// decimation in time, fully unrolled).  With HALF only outputs 0..R/2-1 are
// formed: the last butterfly layer skips its difference half.  With PRE1 the
// caller has done the first layer, whose butterflies are (x[a], x[a + R/2])
// with a unit twiddle: x[a] holds their sum and x[a + R/2] their difference.
template <int R, bool HALF = false, bool PRE1 = false>
__device__ __forceinline__ void dft(float2 (&x)[R])
{
    constexpr int BITS = clog2(R);
    float2 y[R];
#pragma unroll
    for (int i = 0; i < R; i++) y[cbrev(i, BITS)] = x[i];
#pragma unroll
    for (int len = PRE1 ? 4 : 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; j++)
                bfly_tw(y[i + j], y[i + j + len / 2], j * (32 / len), HALF && len == R);
        }
    }
#pragma unroll
    for (int i = 0; i < R; i++) x[i] = y[i];
}

// Tangent form of a twiddle w = f (1 + i t): (t, f) with f = Re w and
// t = Im w / f, so that w v = f (v.x - t v.y, v.y + t v.x) and u +- w v is
// two FMAs for the rotation and one per output component, 6 instructions
// instead of a complex product and two sums (8); the FMA butterfly of Linzer
// and Feig (1993) with a twiddle known only at run time.  t is formed with
// the hardware reciprocal (1 ulp).  The rounding error stays within a few ulp
// of |v| for any angle (a large t comes with a small f); |Re w| < 2^-30 is
// taken as 2^-30, which changes w v by at most 2^-30 |v.x|.
__device__ __forceinline__ float2 lf_form(float2 w)
{
    const float f = fabsf(w.x) >= 0x1p-30f ? w.x : (w.x < 0.0f ? -0x1p-30f : 0x1p-30f);
    return make_float2(w.y * __builtin_amdgcn_rcpf(f), f);
}

// x[r] *= w(r) for r = 1..R-1, then the DFT, with the twiddles of the first
// butterfly layer's second operands fused into it: w(r) is a complex twiddle
// for r < R/2 and the tangent form (lf_form) of one for r >= R/2.  Per pair
// (x[a], x[a + R/2]) a complex product and 6 instructions instead of two
// complex products and a butterfly (16 instructions fewer per radix-16 DFT).
template <int R, class WF>
__device__ __forceinline__ void dft_twin(float2 (&x)[R], WF w)
{
#pragma unroll
    for (int a = 0; a < R / 2; a++) {
        const float2 u = a == 0 ? x[0] : cmul(x[a], w(a));
        const float2 v = x[a + R / 2], tf = w(a + R / 2);
        const float p = __builtin_fmaf(-tf.x, v.y, v.x), q = __builtin_fmaf(tf.x, v.x, v.y);
        x[a] = make_float2(__builtin_fmaf(tf.y, p, u.x), __builtin_fmaf(tf.y, q, u.y));
        x[a + R / 2] = make_float2(__builtin_fmaf(-tf.y, p, u.x), __builtin_fmaf(-tf.y, q, u.y));
    }
    dft<R, false, true>(x);
}

// Entry e of a pass-1 twiddle table (rows of R1 + 1, row k, column r):
// W^(k r) of the R0 R1 points, read from the F-point table tw; with TAN the
// columns r >= R1/2 in tangent form, as dft_twin applies them.
template <int R0, int R1, bool TAN>
__device__ __forceinline__ float2 tw1_entry(const float2* __restrict__ tw, int e, int F)
{
    const int r = e % (R1 + 1);
    const float2 w = tw[((e / (R1 + 1)) * r) % (R0 * R1) * (F / (R0 * R1))];
    return TAN && r >= R1 / 2 ? lf_form(w) : w;
}

// Factors of the twiddles W^(r*e), r < R, W = exp(-2 pi i / F) from the
// table tw: each twiddle is one product of two loaded factors, W^(4ae) *
// W^(be) for r = 4a + b, so a radix-R butterfly loads R/4 + 3 table entries.
template <int R> struct TwFac {
    float2 lo[4], hi[(R / 4) > 0 ? R / 4 : 1];
};

template <int R>
__device__ __forceinline__ void twiddle_fac(TwFac<R>& f, const float2* __restrict__ tw, int e)
{
    f.lo[0] = make_float2(1.0f, 0.0f);
#pragma unroll
    for (int b = 1; b < 4 && b < R; b++) f.lo[b] = tw[b * e];
    f.hi[0] = make_float2(1.0f, 0.0f);
#pragma unroll
    for (int a = 1; a < R / 4; a++) f.hi[a] = tw[4 * a * e];
}

// x[r] *= W^(r*e) for r = 1..R-1 from the factors
template <int R>
__device__ __forceinline__ void twiddle_apply(float2 (&x)[R], const TwFac<R>& f)
{
#pragma unroll
    for (int r = 1; r < R; r++) {
        const int a = r >> 2, b = r & 3;
        float2 w;
        if (a == 0) w = f.lo[b];
        else if (b == 0) w = f.hi[a];
        else w = cmul(f.hi[a], f.lo[b]);
        x[r] = cmul(x[r], w);
    }
}

// x[r] *= W^(r*e) for r = 1..R-1
template <int R>
__device__ __forceinline__ void twiddle_pow(float2 (&x)[R], const float2* __restrict__ tw, int e)
{
    TwFac<R> f;
    twiddle_fac<R>(f, tw, e);
    twiddle_apply<R>(x, f);
}

// --------------------------------------------------------------------------
// Radix plans for F = 2^LG.  Pass 0 is fused with unpack and fringe
// rotation, the last pass forms only channels < nchan and applies the
// fractional-sample rotation.  A block transforms one polarisation and has
// NT = F/R_max threads, so a radix-16 pass hands each thread exactly one
// butterfly; at F = 8192 two blocks (77-79 KB of LDS each) share a CU, so one
// block's barriers and window loads overlap the other's arithmetic.

// Plans V = 0 .. nplans(LG) - 1; the autotuner picks per context
// (fx_autotune).
template <int LG, int V = 0> struct FPlan;
template <> struct FPlan<9, 0>  { static constexpr int NP = 3, R0 = 8,  R1 = 8,  R2 = 8,  R3 = 1, PADB = 4, KIND = 0; };
template <> struct FPlan<10, 0> { static constexpr int NP = 3, R0 = 16, R1 = 8,  R2 = 8,  R3 = 1, PADB = 4, KIND = 0; };
template <> struct FPlan<10, 1> { static constexpr int NP = 3, R0 = 8,  R1 = 8,  R2 = 16, R3 = 1, PADB = 4, KIND = 0; };
template <> struct FPlan<11, 0> { static constexpr int NP = 3, R0 = 16, R1 = 16, R2 = 8,  R3 = 1, PADB = 4, KIND = 0; };
template <> struct FPlan<11, 1> { static constexpr int NP = 3, R0 = 8,  R1 = 16, R2 = 16, R3 = 1, PADB = 4, KIND = 0; };
template <> struct FPlan<12, 0> { static constexpr int NP = 3, R0 = 16, R1 = 16, R2 = 16, R3 = 1, PADB = 5, KIND = 0; };
template <> struct FPlan<12, 1> { static constexpr int NP = 4, R0 = 8,  R1 = 8,  R2 = 8,  R3 = 8, PADB = 5, KIND = 0; };
template <> struct FPlan<13, 0> { static constexpr int NP = 4, R0 = 16, R1 = 16, R2 = 16, R3 = 2, PADB = 5, KIND = 0; };
template <> struct FPlan<13, 1> { static constexpr int NP = 4, R0 = 8,  R1 = 16, R2 = 16, R3 = 4, PADB = 4, KIND = 0; };
template <> struct FPlan<14, 0> { static constexpr int NP = 4, R0 = 16, R1 = 16, R2 = 16, R3 = 4, PADB = 5, KIND = 0; };
template <> struct FPlan<14, 1> { static constexpr int NP = 4, R0 = 16, R1 = 16, R2 = 8,  R3 = 8, PADB = 5, KIND = 0; };
// PADB: LDS skew of one point per 2^PADB points (pad), the one of 16 or 32
// that measured fewer LDS bank-conflict cycles for the plan's access pattern
// (the LDS has 64 banks of 4 bytes; ds_read2/write2 pairs and single 8-byte
// accesses conflict differently, so no one skew suits every plan).
// KIND 0: Stockham passes; KIND 1: decimation in frequency with in-place
// passes (dif_pass); KIND 2: Stockham with the last radix-2 pass done across
// the two half-waves (v_permlane32_swap) instead of through LDS.
template <> struct FPlan<12, 2> { static constexpr int NP = 3, R0 = 16, R1 = 16, R2 = 16, R3 = 1, PADB = 4, KIND = 1; };
template <> struct FPlan<13, 2> { static constexpr int NP = 4, R0 = 16, R1 = 16, R2 = 16, R3 = 2, PADB = 4, KIND = 2; };
template <> struct FPlan<14, 2> { static constexpr int NP = 4, R0 = 16, R1 = 16, R2 = 16, R3 = 4, PADB = 4, KIND = 1; };
// Three passes with 512 threads (real input only): one LDS exchange fewer than
// the four-pass plans, and two waves per SIMD leave 256 registers a lane for
// the radix-32 butterflies and the output chain.
template <> struct FPlan<14, 3> { static constexpr int NP = 3, R0 = 16, R1 = 32, R2 = 32, R3 = 1, PADB = 5, KIND = 0; };

// number of plans per FFT length
__host__ __device__ constexpr int nplans(int lg) { return lg == 9 ? 1 : lg == 14 ? 4 : (lg >= 12 && lg <= 13) ? 3 : 2; }

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// CPLX: complex input, F = nchan and every bin a channel; otherwise real
// input, nchan = F/2.  stride = 2^floor(log2(nchan)/2) either way.
template <int LG, int V = 0, bool CPLX = false> struct Geo {
    using PL_ = FPlan<LG, V>;
    static constexpr int LGV = LG;
    static constexpr int KIND = PL_::KIND;
    static constexpr int F = 1 << LG;
    static constexpr int NCH = CPLX ? F : F / 2;
    static constexpr int STRIDE = 1 << ((CPLX ? LG : LG - 1) / 2);
    static constexpr int SUB = 2 * STRIDE;
    static constexpr int NSTEP = F / SUB;
    static constexpr int NFSTEP = NCH / STRIDE;
    static constexpr int NTAB = SUB + NSTEP + STRIDE + NFSTEP;
    static constexpr int RTAB = NTAB + NSTEP;                // table stride in memory (fx_rtab_len)
    static constexpr int NP = PL_::NP, R0 = PL_::R0, R1 = PL_::R1, R2 = PL_::R2, R3 = PL_::R3;
    static constexpr int RMAX = cmax(cmax(R0, R1), cmax(R2, R3));
    static constexpr int NT = F / RMAX;
    static constexpr int RL = NP == 4 ? R3 : R2;            // last radix
    static constexpr int PL = F / RL;                       // transform length before the last pass
    static constexpr int NOUT = NCH / NT;                   // channels per thread per FFT
    static constexpr int PADB = PL_::PADB;
    static constexpr int PM = 1 << PADB;                    // points per skew step
    static constexpr int FP = F + (F >> PADB);              // padded LDS row
    // LDS index of point n.  pad(b + s) = pad(b) + pad(s) whenever
    // (b mod PM) + (s mod PM) < PM; the butterfly addressing meets that with
    // s = r*S for S a multiple of PM, or for S dividing PM and b mod PM < S,
    // so every offset from a butterfly's base address is a constant.
    static __host__ __device__ constexpr int pad(int n) { return n + (n >> PADB); }
    // waves per SIMD: at F = 8192 the LDS allows two 8-wave blocks per CU, at
    // F = 16384 one block of 16 (or, radix 32, 8) waves (at F = 4096 see fstage_wpe)
    static constexpr int WPE = LG == 14 ? NT / 256 : LG >= 13 ? 4 : 3;
    static_assert(R0 * R1 * R2 * R3 == F, "plan must factor F");
    static_assert((F / R0) % SUB == 0, "pass-0 rotator reuse needs F/R0 to be a multiple of sub");
    static_assert(R0 == 16 || (R0 == 8 && F / R0 >= 64), "pass-0 writes pad(i*R0 + r) = pad(i*R0) + r");
    static_assert(NT % PM == 0, "last-pass reads pad(i + m*NT + r*TL) = pad(i) + pad(m*NT) + pad(r*TL)");
    static_assert(PM % R0 == 0, "pass-0 writes pad(i*R0 + r) = pad(i*R0) + r");
    static_assert(NT <= 1024, "block too large");
};


// Each byte of w holds time slots t = 0, 1 of both pols, pol p of slot t at
// bits 2*(2t+p) (kernel.c:195-200).  Returns, per byte, the two codes of pol
// p packed as slot0 | slot1 << 2.
__device__ __forceinline__ uint32_t pol_codes(uint32_t w, int pol)
{
    const uint32_t t = w >> (2 * pol);
    return (t & 0x03030303u) | ((t >> 2) & 0x0c0c0c0cu);
}

// Load of data that no kernel writes while this one runs, through the
// constant address space: at a wave-uniform address it is a scalar load.
__device__ __forceinline__ float2 load_const(const float2* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const float cfloat;
    const cfloat* q = (const cfloat*)p;
    return make_float2(q[0], q[1]);
#else
    return *p;
#endif
}

// (offset, valid) of a delay-table entry through the constant address space
// (scalar loads at a block-uniform address).
__device__ __forceinline__ int2 load_delay_ov(const FxDelay* d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const int cint;
    const cint* q = (const cint*)&d->offset;
    return make_int2(q[0], q[1]);
#else
    return make_int2(d->offset, d->valid);
#endif
}

// Per-thread storage of the G spectra of the channels a thread owns.  HALF
// (FP8 only): the chain holds G/2 FFTs, and each half of a row is stored
// when its G/2 FFTs are done.
template <int PREC, int G, int NOUT, bool HALF = false> struct OutRegs;
// The G FFTs of a group pass through a register shift chain: after each FFT
// the chain moves down one slot and the new spectrum enters slot G-1, so the
// FFT loop can stay rolled while every register index is static.
template <int G, int NOUT, bool HALF> struct OutRegs<FX_PREC_F32, G, NOUT, HALF> {
    static_assert(!HALF, "half rows: FP8 only");
    float2 v[G][NOUT];
    __device__ __forceinline__ void shift(int) {
#pragma unroll
        for (int g = 0; g + 1 < G; g++)
#pragma unroll
            for (int u = 0; u < NOUT; u++) v[g][u] = v[g + 1][u];
    }
    __device__ __forceinline__ void zero_last(int) {
#pragma unroll
        for (int u = 0; u < NOUT; u++) v[G - 1][u] = make_float2(0.0f, 0.0f);
    }
    __device__ __forceinline__ void put_last(int u, float2 x, float, int) { v[G - 1][u] = x; }
};
template <int G, int NOUT, bool HALF> struct OutRegs<FX_PREC_F16, G, NOUT, HALF> {
    static_assert(!HALF, "half rows: FP8 only");
    uint32_t v[G][NOUT];
    __device__ __forceinline__ void shift(int) {
#pragma unroll
        for (int g = 0; g + 1 < G; g++)
#pragma unroll
            for (int u = 0; u < NOUT; u++) v[g][u] = v[g + 1][u];
    }
    __device__ __forceinline__ void zero_last(int) {
#pragma unroll
        for (int u = 0; u < NOUT; u++) v[G - 1][u] = 0u;
    }
    __device__ __forceinline__ void put_last(int u, float2 x, float, int) {
        const __half2 h = __floats2half2_rn(x.x, x.y);
        v[G - 1][u] = *reinterpret_cast<const uint32_t*>(&h);
    }
};
// FP8: a word holds two FFTs (re,im e4m3 each) as 16-bit halves, slot order
// (low, high) = (earlier, later).  FFT gi of the group enters half gi & 1 of
// the last word, and before each even FFT the words move down by one: one
// word move per output and FFT pair instead of a half-word shift per FFT.
// With HALF the chain is W = G/4 words (a half row), emptied half-way.
template <int G, int NOUT, bool HALF> struct OutRegs<FX_PREC_F8, G, NOUT, HALF> {
    static constexpr int W = HALF ? G / 4 : G / 2;
    uint32_t v[W][NOUT];
    __device__ __forceinline__ void shift(int gi) {
        if (gi & 1) return;
#pragma unroll
        for (int u = 0; u < NOUT; u++)
#pragma unroll
            for (int w = 0; w + 1 < W; w++) v[w][u] = v[w + 1][u];
    }
    __device__ __forceinline__ void zero_last(int gi) {
#pragma unroll
        for (int u = 0; u < NOUT; u++) v[W - 1][u] = (gi & 1) ? (v[W - 1][u] & 0xffffu) : 0u;
    }
    __device__ __forceinline__ void put_last(int u, float2 x, float si, int gi) {
        const float a = fminf(fmaxf(x.x * si, -448.0f), 448.0f);
        const float b = fminf(fmaxf(x.y * si, -448.0f), 448.0f);
        if (gi & 1)
            v[W - 1][u] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, (int)v[W - 1][u], true);
        else
            v[W - 1][u] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    }
};

template <int PREC> struct GroupSize;
template <> struct GroupSize<FX_PREC_F32> { static constexpr int G = 2; };
template <> struct GroupSize<FX_PREC_F16> { static constexpr int G = 4; };
template <> struct GroupSize<FX_PREC_F8>  { static constexpr int G = 8; };

int fx_fstage_group_size(int prec)
{
    return prec == FX_PREC_F8 ? GroupSize<FX_PREC_F8>::G
         : prec == FX_PREC_F16 ? GroupSize<FX_PREC_F16>::G : GroupSize<FX_PREC_F32>::G;
}

// --------------------------------------------------------------------------
// Middle pass: radix R on sub-transforms of length P (Stockham, in place in
// LDS with a barrier between the reads and the writes).

// TWS is the stride of the twiddle table: 1 when tw holds exp(-2 pi i k / 2^LG),
// 2 when a 2^LG-point sub-transform reads the table of a 2^(LG+1)-point one.
// fac: when not null, the thread's twiddle factors of this pass (one per
// butterfly m), loaded before the FFT loop, in place of the table tw.
// TAN: the LDS table tw1 holds tangent forms in its columns r >= R/2
// (tw1_entry), applied in the DFT's first butterfly layer (dft_twin).
template <class Gm, int R, int P, bool TW_LDS, int TWS = 1, bool TAN = false>
__device__ __forceinline__ void mid_pass(float2* buf, const float2* tw1, const float2* __restrict__ tw, int tid,
                                         const TwFac<R>* fac = nullptr)
{
    constexpr int F = Gm::F, NT = Gm::NT, FP = Gm::FP;
    constexpr int T = F / R, PER = T / NT;
    static_assert(PER * NT == T, "plan/thread mismatch");
    constexpr int PM = Gm::PM;
    static_assert(T % PM == 0, "padded addressing needs T to be a multiple of PM");
    static_assert(P % PM == 0 || (PM % P == 0 && (P * R) % PM == 0), "write offsets pad(r*P)");
    float2 x[PER][R];
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int i = tid + m * NT;
        const int k = i & (P - 1);
        // T is a multiple of PM, so pad(i + r*T) = pad(i) + pad(r*T): one base
        // address and immediate offsets per butterfly
        const float2* rd = buf + Gm::pad(i);
#pragma unroll
        for (int r = 0; r < R; r++) x[m][r] = rd[Gm::pad(r * T)];
        if (TW_LDS) {
            // tw1 is k-major with row stride R + 1: a per-lane base and an
            // immediate offset per r, and 16 lanes hit 16 distinct banks
            const float2* tk = tw1 + k * (R + 1);
            if constexpr (TAN) {
                dft_twin<R>(x[m], [&](int r) { return tk[r]; });
            } else {
#pragma unroll
                for (int r = 1; r < R; r++) x[m][r] = cmul(x[m][r], tk[r]);
                dft<R>(x[m]);
            }
        } else {
            if (fac) twiddle_apply<R>(x[m], fac[m]);
            else twiddle_pow<R>(x[m], tw, k * (TWS * F / (P * R)));
            dft<R>(x[m]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int i = tid + m * NT;
        const int k = i & (P - 1);
        const int j = (i - k) * R + k;
        // j mod PM = k < P (or P a multiple of PM): pad(j + r*P) = pad(j) + pad(r*P)
        float2* wr = buf + Gm::pad(j);
#pragma unroll
        for (int r = 0; r < R; r++) wr[Gm::pad(r * P)] = x[m][r];
    }
    __syncthreads();
}

// --------------------------------------------------------------------------
// Decimation-in-frequency pass, in place: blocks of L points, butterflies on
// the R points j + r*S (S = L/R) of a block, then the twiddle W_L^(j*k) on
// output k, written back to the points the thread read.  No thread touches
// another's points within a pass, so a pass needs one barrier (after it)
// instead of the two of a Stockham pass.  After the passes the point at
// position sum_p k_p*S_p holds frequency sum_p k_p*(R_0*...*R_(p-1)).

template <class Gm, int R, int L>
__device__ __forceinline__ void dif_pass(float2* buf, const float2* __restrict__ tw, int tid)
{
    constexpr int F = Gm::F, NT = Gm::NT;
    constexpr int S = L / R;
    constexpr int T = F / R, PER = T / NT;
    static_assert(PER * NT == T, "plan/thread mismatch");
    static_assert(S > 1, "the last pass has its own code");
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int q = tid + m * NT;
        const int j = q & (S - 1);
        const int base = (q - j) * R + j;         // (q / S) * L + j
        float2 x[R];
        if constexpr (S % Gm::PM == 0 || (Gm::PM % S == 0 && L % Gm::PM == 0)) {
            // base mod PM = j < S (or S a multiple of PM): pad(base + r*S) = pad(base) + pad(r*S)
            float2* p = buf + Gm::pad(base);
#pragma unroll
            for (int r = 0; r < R; r++) x[r] = p[Gm::pad(r * S)];
            dft<R>(x);
            twiddle_pow<R>(x, tw, j * (F / L));
#pragma unroll
            for (int r = 0; r < R; r++) p[Gm::pad(r * S)] = x[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) x[r] = buf[Gm::pad(base + r * S)];
            dft<R>(x);
            twiddle_pow<R>(x, tw, j * (F / L));
#pragma unroll
            for (int r = 0; r < R; r++) buf[Gm::pad(base + r * S)] = x[r];
        }
    }
    __syncthreads();
}

// Frequency of the last DIF pass's butterfly q (q < F/R_last), output k = 0:
// the digits of q in radices (R0, R1[, R2]), most significant first, reversed.
template <class Gm>
__device__ __forceinline__ int dif_freq(int q)
{
    constexpr int R0 = Gm::R0, R1 = Gm::R1;
    if constexpr (Gm::NP == 4) {
        constexpr int R2 = Gm::R2;
        const int k2 = q % R2, k1 = (q / R2) % R1, k0 = q / (R1 * R2);
        return k0 + R0 * (k1 + R1 * k2);
    } else {
        const int k1 = q % R1, k0 = q / R1;
        return k0 + R0 * k1;
    }
}

// Channel of a thread's output m*RO + r (RO outputs per last-pass butterfly).
template <class Gm>
__device__ __forceinline__ int out_chan(int tid, int m, int r)
{
    if constexpr (Gm::KIND == 2)      // cross-lane plan: channel k + 256 (m + 8 up) (+ F/2 for r = 1)
        return ((tid >> 6) * 32 + (tid & 31)) + 256 * (m + 8 * ((tid >> 5) & 1)) + r * (Gm::F / 2);
    const int i = tid + m * Gm::NT;
    return (Gm::KIND == 1 ? dif_freq<Gm>(i) : i) + r * Gm::PL;
}

// --------------------------------------------------------------------------
// The fused F-stage kernel: one block = one antenna, one polarisation, G
// consecutive FFTs.

// Waves per SIMD the kernel is compiled for.  At F = 4096 the LDS admits four
// 4-wave blocks per CU, which needs at most 128 VGPRs: real input with FP16
// or FP32 output fits, FP8 output and complex input (twice the channels) do
// not and keep three.
template <int LG, int V, int PREC, bool CPLX>
__host__ __device__ constexpr int fstage_wpe()
{
    return (LG == 12 && !CPLX && PREC != FX_PREC_F8) ? 4 : Geo<LG, V, CPLX>::WPE;
}

// Tangent-form twiddles (lf_form) in the first two passes: fewer VALU
// operations per point, but more live registers.  Below F = 8192 every plan
// that uses them (Stockham: pass 1 from the LDS table) has room for them and
// gains 1-2.5% of its VALU instructions; at F >= 8192 the kernels are compiled
// for the 128-register budget of four waves per SIMD, where the forms brought
// scratch spills and the 8192-point FP16 F-stage measured 3-4% slower; the
// 4096-point FP8 kernel of plan 0 also spills with them.
template <int LG, int V, int PREC, bool CPLX>
__host__ __device__ constexpr bool fstage_tangent()
{
    return LG <= 12 && FPlan<LG, V>::KIND != 1 && !(LG == 12 && V == 0 && PREC == FX_PREC_F8);
}

// Pass 0 of real input reads its fringe steps with scalar loads when the
// step index is uniform across each wave; with one pass-0 butterfly per
// thread it reads them from the tables' second copy of the steps, as one
// contiguous block.
template <int LG, int V, bool CPLX>
__host__ __device__ constexpr bool fstage_sstep()
{
    using Gm = Geo<LG, V, CPLX>;
    return !CPLX && (Gm::SUB % 64 == 0) && (Gm::NT % Gm::SUB == 0);
}
template <int LG, int V, bool CPLX>
__host__ __device__ constexpr bool fstage_sperm()
{
    using Gm = Geo<LG, V, CPLX>;
    return fstage_sstep<LG, V, CPLX>() && Gm::F / Gm::R0 == Gm::NT;
}

template <int LG, bool CPLX>
static int step_q_lg(int plan)
{
    using G0 = Geo<LG, 0, CPLX>;
    if constexpr (nplans(LG) > 3 && !CPLX)
        if (plan == 3) return fstage_sperm<LG, 3, CPLX>() ? G0::F / Geo<LG, 3, CPLX>::R0 / G0::SUB : 0;
    if constexpr (nplans(LG) > 2)
        if (plan == 2) return fstage_sperm<LG, 2, CPLX>() ? G0::F / Geo<LG, 2, CPLX>::R0 / G0::SUB : 0;
    if constexpr (nplans(LG) > 1)
        if (plan == 1) return fstage_sperm<LG, 1, CPLX>() ? G0::F / Geo<LG, 1, CPLX>::R0 / G0::SUB : 0;
    return fstage_sperm<LG, 0, CPLX>() ? G0::F / G0::R0 / G0::SUB : 0;
}

int fstage_step_q(int lg, int plan, int iscomplex)
{
    if (iscomplex) return 0;
    switch (lg) {
    case 9:  return step_q_lg<9, false>(plan);
    case 10: return step_q_lg<10, false>(plan);
    case 11: return step_q_lg<11, false>(plan);
    case 12: return step_q_lg<12, false>(plan);
    case 13: return step_q_lg<13, false>(plan);
    case 14: return step_q_lg<14, false>(plan);
    default: return 0;                        // the 32768-point halves read the steps from the LDS
    }
}

// Work of F-stage workgroup `lid` (numbered after the XCD-aware reordering):
// tiles of `tile` antennas outermost, then the launch's FFT groups, then the
// tile's signals.  The eight workgroups whose rows share a 128-byte line of
// the X input (four antennas, two polarisations; tile is a multiple of four)
// run side by side, the two polarisations of an antenna (same samples, same
// rotator tables) are neighbours, and one signal's groups follow each other
// closely, so its FP8 inverse scales stay in the L2 (a group-major order
// re-read them from HBM for every group).  The last tile holds nant % tile
// antennas.
__device__ __forceinline__ void fstage_work(int lid, int nant, int ngroups, int tile, int& grp, int& ant, int& pol)
{
    const int full = nant / tile;
    const int per_tile = 2 * tile * ngroups;
    int t, r, w;
    if (lid < full * per_tile) {
        t = lid / per_tile;
        r = lid - t * per_tile;
        w = 2 * tile;
    } else {
        t = full;
        r = lid - full * per_tile;
        w = 2 * (nant - tile * full);
    }
    grp = r / w;
    const int sl = r - grp * w;
    ant = tile * t + (sl >> 1);
    pol = sl & 1;
}

// Antennas per work tile: enough that at most eight FFT groups are in flight
// on an XCD (32 CUs x the kernel's resident workgroups per CU), and at least
// four.  More groups in flight put more lines 2-8 KB apart (one per group and
// channel) into the L2 at once, and the F-stage's writes grew from 1.04x to
// 1.14x the spectra at 256 antennas x 1024 channels FP8 (six workgroups per
// CU: 16 antennas per tile); at 64 x 4096 (two per CU) four antennas measured
// fastest.  The occupancy query runs once per kernel, outside stream capture;
// a launch captured before it uses four.
static constexpr int kMinTileAnt = 4;
static int fstage_tile_ant(std::atomic<int>& cache, const void* fn, int nthreads, hipStream_t s)
{
    int t = cache.load(std::memory_order_relaxed);
    if (t > 0) return t;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return kMinTileAnt;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nthreads, 0) != hipSuccess || nb < 1) nb = 1;
    t = kMinTileAnt;
    while (32 * nb > 16 * t) t *= 2;
    cache.store(t, std::memory_order_relaxed);
    return t;
}

template <int LG, int V, int PREC, bool TAP, bool CPLX = false>
__global__ __launch_bounds__((Geo<LG, V>::NT)) __attribute__((amdgpu_waves_per_eu(fstage_wpe<LG, V, PREC, CPLX>())))
void fx_fstage_kernel(FStageArgs args)
{
    using Gm = Geo<LG, V, CPLX>;
    constexpr bool DIF = (Gm::KIND == 1);
    constexpr bool XL = (Gm::KIND == 2);
    constexpr int F = Gm::F, NCH = Gm::NCH, NT = Gm::NT, FP = Gm::FP;
    constexpr int STRIDE = Gm::STRIDE, SUB = Gm::SUB, NSTEP = Gm::NSTEP, NFSTEP = Gm::NFSTEP;
    constexpr int R0 = Gm::R0, R1 = Gm::R1, R2 = Gm::R2, RL = Gm::RL, PL = Gm::PL;
    constexpr int G = GroupSize<PREC>::G;
    constexpr int T0 = F / R0, PER0 = T0 / NT;
    constexpr int TL = F / RL, PERL = TL / NT;
    static_assert(PER0 * NT == T0 && PERL * NT == TL, "plan/thread mismatch");
    constexpr int RO = CPLX ? RL : RL / 2;               // channels per last-pass butterfly
    constexpr int NOUT = PERL * RO;
    static_assert(NOUT == Gm::NOUT, "output count");
    constexpr int NTW1 = R0 * R1;

    __shared__ __align__(16) float2 buf[FP];
    // The smaller tables (at F = 4096 in one LDS array: each __shared__ array
    // is rounded up to 16 bytes, and the 16 bytes saved let four blocks share
    // a CU; elsewhere as separate arrays, which measured faster at F = 8192):
    //  * tw1[k][r] = exp(-2 pi i k r / (R0 R1)), rows R1 + 1 apart, the last
    //    row's spare entry left out;
    //  * tab: the rotator tables of this (FFT, antenna);
    //  * bytes (real input): the FFT window with each byte recoded to this
    //    pol's two codes (pol_codes); F samples span at most F/2 + 4 bytes
    //    from the window's aligned start, so F/8 + 1 words;
    //  * lut (real input): the level table lut[2*code pair + time slot]; its
    //    32 words sit in 32 banks, so lanes that read one entry share a
    //    broadcast and lanes that read different entries never conflict.
    // Complex input reads its int8 pairs from global memory.
    constexpr int NTW1L = R0 * (R1 + 1) - 1;
    constexpr int NWIN = CPLX ? 4 : F / 8 + 1;
    constexpr int NLUT = CPLX ? 2 : 32;
    constexpr bool MERGE = LG == 12;
    __shared__ __align__(16) float2 aux[MERGE ? NTW1L + Gm::NTAB + (NWIN + NLUT + 1) / 2 : 1];
    __shared__ __align__(16) float2 tw1_s[MERGE ? 1 : NTW1L];
    __shared__ __align__(16) float2 tab_s[MERGE ? 1 : Gm::NTAB];
    __shared__ __align__(16) uint32_t bytes_s[MERGE ? 1 : NWIN];
    __shared__ __align__(16) float lut_s[MERGE ? 1 : NLUT];
    float2* const tw1 = MERGE ? aux : tw1_s;
    float2* const tab = MERGE ? aux + NTW1L : tab_s;
    uint32_t* const bytes = MERGE ? reinterpret_cast<uint32_t*>(aux + NTW1L + Gm::NTAB) : bytes_s;
    float* const lut = MERGE ? reinterpret_cast<float*>(aux + NTW1L + Gm::NTAB) + NWIN : lut_s;
    const float2* rsub = tab;
    const float2* rstep = tab + SUB;
    const float2* frsub = tab + SUB + NSTEP;
    const float2* frstep = tab + SUB + NSTEP + STRIDE;
    constexpr int TPT = (Gm::NTAB + NT - 1) / NT;       // table entries per thread
    constexpr int WPT = CPLX ? 1 : (F / 8 + 1 + NT - 1) / NT;  // window words per thread (real input)

    const FxPlan& p = args.p;

    // XCD-aware order: hardware places block b on XCD b % 8; give each XCD a
    // contiguous range of work (fstage_work) so the G-FFT runs written by
    // neighbouring antennas meet in the same L2.
    const int nblocks = gridDim.x;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int lid = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
    int grp, ant, pol;
    fstage_work(lid, p.nant, args.ngroups, args.tile_ant, grp, ant, pol);
    const int fgo = args.fg0 + grp;
    const int sig = 2 * ant + pol;
    // X input row of (channel 0, this group, this signal); channel c adds c * fx_xin_pitch rows
    const int kq = fgo / args.kc;
    const long long cpitch = fx_xin_pitch(args.kc, p.Spad);
    const long long xrow0 = (long long)kq * NCH * cpitch + (long long)(fgo - kq * args.kc) * p.Spad + sig;
    int tid = threadIdx.x;

    constexpr bool TAN = fstage_tangent<LG, V, PREC, CPLX>();
    for (int e = tid; e < NTW1L; e += NT) tw1[e] = tw1_entry<R0, R1, TAN>(args.tw, e, F);
    if constexpr (!CPLX) {
        if (tid < 32) {
            // code pair e = slot-0 code | slot-1 code << 2; levels of kernel.c:191-192
            const float hi = 3.3359f;
            const float lv[4] = {-hi, -1.0f, 1.0f, hi};
            lut[tid] = lv[((tid >> 1) >> (2 * (tid & 1))) & 3];
        }
    }

    const uint8_t* src = args.data + (long long)ant * args.data_stride;
    // FP8 with complex input at F >= 8192 (16 channels per thread): half rows,
    // so the output chain takes half the registers (the kernels spilled 72-148
    // bytes per lane holding whole rows)
    constexpr bool F8HALF = PREC == FX_PREC_F8 && CPLX && NOUT >= 16;
    OutRegs<PREC, G, NOUT, F8HALF> outr;
    // half h of every output row: 8 bytes, two words of the chain
    [[maybe_unused]] auto store_half = [&](int h, int t) {
        if constexpr (F8HALF) {
#pragma unroll
            for (int m = 0; m < PERL; m++)
#pragma unroll
                for (int r = 0; r < RO; r++) {
                    const int ou = m * RO + r;
                    const long long row = xrow0 + (long long)out_chan<Gm>(t, m, r) * cpitch;
                    reinterpret_cast<uint2*>(args.out)[row * 2 + h] = make_uint2(outr.v[0][ou], outr.v[1][ou]);
                }
        }
    };

    // The next FFT's rotator tables and packed-byte window are loaded into
    // registers while the current FFT runs; the delay entry (offset, valid)
    // that their addresses need is loaded one FFT earlier still, so no FFT
    // waits for a load issued after its predecessor's barrier.
    struct Prefetch {
        int valid, offset;
        float2 t[TPT];
        uint32_t w[WPT];
    };
    auto dload = [&](int gi) { return load_delay_ov(args.delays + (long long)(grp * G + gi) * p.nant + ant); };
    auto fetch = [&](int gi, Prefetch& f, int2 d) {
        const long long fa = (long long)(grp * G + gi) * p.nant + ant;
        f.offset = d.x;
        f.valid = d.y;
        if (!f.valid) return;
        const float2* tb = args.rtab + fa * Gm::RTAB;
#pragma unroll
        for (int t = 0; t < TPT; t++) {
            const int e = threadIdx.x + t * NT;
            f.t[t] = e < Gm::NTAB ? tb[e] : make_float2(0.0f, 0.0f);
        }
        if constexpr (!CPLX) {
            const int bb = (f.offset >> 1) & ~3;
            const int nd = ((((f.offset + F - 1) >> 1) - bb) >> 2) + 1;
            const uint32_t* s32 = reinterpret_cast<const uint32_t*>(src + bb);
#pragma unroll
            for (int w = 0; w < WPT; w++) {
                const int k = threadIdx.x + w * NT;
                f.w[w] = k < nd ? s32[k] : 0u;
            }
        }
    };
    Prefetch cur;
    fetch(0, cur, dload(0));
    int2 dnext = dload(G > 1 ? 1 : 0);

    // XL: lanes 0-31 of wave w own the third-pass butterflies k = 32w + lane,
    // lanes 32-63 the butterflies 256 + k; their outputs pair up in the last
    // radix-2 step X[c] = A[c] + W^c B[c], c = k + 256*rho.  W^(k + 256*rho) =
    // W^k (-j)^up W_32^r with rho = r + 8*up, so W^k (-j)^up is fixed per lane.
    [[maybe_unused]] float2 xl_wk = make_float2(1.0f, 0.0f);
    if constexpr (XL) {
        static_assert(LG == 13 && NT == 512 && R0 == 16 && R1 == 16 && R2 == 16 && RL == 2 && STRIDE == 64,
                      "cross-lane last pass laid out for 16*16*16*2");
        const int up = (threadIdx.x >> 5) & 1;
        const int k = (threadIdx.x >> 6) * 32 + (threadIdx.x & 31);
        const float2 w = args.tw[k];
        xl_wk = up ? make_float2(w.y, -w.x) : w;
    }

    // XL at G = 4 (FP16; the FFT loop fully unrolled leaves the registers): the
    // third pass's fifteen twiddles and the last step's eight, formed once per
    // block (at G = 2, FP32, forming them for two FFTs measured slower)
    constexpr bool XLH = XL && G == 4 && !CPLX;
    [[maybe_unused]] float2 xw3[XLH ? 16 : 1], xwl[XLH ? 8 : 1];
    if constexpr (XLH) {
        TwFac<16> f3;
        twiddle_fac<16>(f3, args.tw, ((threadIdx.x >> 6) * 32 + (threadIdx.x & 31)) * (F / (256 * 16)));
        xw3[0] = make_float2(1.0f, 0.0f);
#pragma unroll
        for (int r = 1; r < 16; r++) {
            const int a = r >> 2, b = r & 3;
            xw3[r] = a == 0 ? f3.lo[b] : b == 0 ? f3.hi[a] : cmul(f3.hi[a], f3.lo[b]);
        }
#pragma unroll
        for (int r = 0; r < 8; r++) xwl[r] = twc(xl_wk, r);
    }
    // FP8: the inverse scales of this signal's output channels, loaded once
    // where the register budget allows (three waves per SIMD; the radix-32
    // 16384-point kernel, at two, spills with them)
    constexpr bool SREG = PREC == FX_PREC_F8 && fstage_wpe<LG, V, PREC, CPLX>() == 3 && !CPLX;
    [[maybe_unused]] float sreg[SREG ? NOUT : 1];
    if constexpr (SREG) {
#pragma unroll
        for (int m = 0; m < PERL; m++)
#pragma unroll
            for (int r = 0; r < RO; r++)
                sreg[m * RO + r] = args.scale_inv[(long long)sig * NCH + out_chan<Gm>(threadIdx.x, m, r)];
    }

    // Twiddle factors whose exponents depend only on the thread, loaded once
    // where the register budget allows (three waves per SIMD, and at F = 4096
    // and, except FP8, F = 16384: 103-127 registers, 1-5% fewer VALU
    // instructions; the FP8 16384-point kernels spill with them, the FP16
    // radix-32 one 36 bytes, and measured 3% faster with them than without):
    // the third pass's of a four-pass Stockham plan and the last pass's.
    constexpr bool TWREG = (fstage_wpe<LG, V, PREC, CPLX>() == 3 || LG == 12 || (LG == 14 && PREC != FX_PREC_F8)) &&
                           !XL && !DIF && !CPLX;
    constexpr int PER2 = (F / R2) / NT > 0 ? (F / R2) / NT : 1;
    [[maybe_unused]] TwFac<R2> fac2[TWREG && Gm::NP == 4 ? PER2 : 1];
    [[maybe_unused]] TwFac<RL> facl[TWREG ? PERL : 1];
    if constexpr (TWREG) {
        if constexpr (Gm::NP == 4) {
#pragma unroll
            for (int m = 0; m < PER2; m++)
                twiddle_fac<R2>(fac2[m], args.tw, ((int)threadIdx.x + m * NT) % (R0 * R1) * (F / (R0 * R1 * R2)));
        }
#pragma unroll
        for (int m = 0; m < PERL; m++) twiddle_fac<RL>(facl[m], args.tw, (int)threadIdx.x + m * NT);
    }

    // The FFT loop is unrolled (G <= 4) or, for FP8 (G = 8, where eight copies
    // measured slower: I-cache), runs two FFTs per trip where the register
    // budget allows (three waves per SIMD or fewer): no register moves of the
    // output chain and the prefetch between the FFTs of one trip.
    constexpr int UNR = G <= 4 ? G : (fstage_wpe<LG, V, PREC, CPLX>() <= 3 && !CPLX) ? 2 : 1;
#pragma unroll UNR
    for (int gi = 0; gi < G; gi++) {
        // Redefine tid for the compiler on every FFT: the LDS addresses derived
        // from it are then formed where they are used instead of being hoisted
        // out of the loop and held in registers across all passes.
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid >= 0 && tid < NT);
        if constexpr (F8HALF) {
            if (gi == G / 2) store_half(0, tid);  // FFTs 0 .. G/2-1 are done
        }
        outr.shift(gi);
        const int fft = grp * G + gi;             // FFT of this launch
        __syncthreads();                          // previous FFT done with LDS
        if (!cur.valid) {                         // uniform across the block
            outr.zero_last(gi);
            if (TAP) {
#pragma unroll
                for (int m = 0; m < PERL; m++)
#pragma unroll
                    for (int r = 0; r < RO; r++)
                        args.tap[(((long long)fft * p.nant + ant) * 2 + pol) * NCH + tid + m * NT + r * PL] =
                            make_float2(0.0f, 0.0f);
            }
            if (gi + 1 < G) {
                fetch(gi + 1, cur, dnext);
                if (gi + 2 < G) dnext = dload(gi + 2);
            }
            continue;
        }
        // ---- tables and window of this FFT from registers to LDS
        const int o = cur.offset;
        const int bbase = (o >> 1) & ~3;
#pragma unroll
        for (int t = 0; t < TPT; t++) {
            const int e = tid + t * NT;
            if (e < Gm::NTAB) tab[e] = cur.t[t];
        }
        if constexpr (!CPLX) {
#pragma unroll
            for (int w = 0; w < WPT; w++) {
                const int k = tid + w * NT;
                if (k < NWIN) bytes[k] = pol_codes(cur.w[w], pol);
            }
        }
        __syncthreads();
        // pass 0 at raised wave priority at F = 8192 with FP16 output (two
        // eight-wave blocks per CU): its dependent LDS reads (window byte, then
        // level) issue ahead of the other block's arithmetic, which fills the
        // VALU meanwhile (-6% F-stage time); measured 3% slower with FP32
        // output (two FFTs per block) and 1% slower with many small blocks per
        // CU (F = 2048, FP8), so those keep one priority
        constexpr bool PRIO = LG == 13 && PREC == FX_PREC_F16;
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
        if (gi + 1 < G) {
            fetch(gi + 1, cur, dnext);
            if (gi + 2 < G) dnext = dload(gi + 2);
        }

        // ---- pass 0: unpack + fringe rotate + radix-R0 butterfly
        if constexpr (CPLX) {
            // sample n = o + i + r*T0 is the word at byte 4n: this pol's int8
            // (re, im) in its low (pol 0) or high (pol 1) half
            const uint32_t* w32 = reinterpret_cast<const uint32_t*>(src) + o;
#pragma unroll
            for (int m = 0; m < PER0; m++) {
                const int i = tid + m * NT;
                const float2 rs = rsub[i & (SUB - 1)];
                const float2* sp = rstep + i / SUB;
                float2 x[R0];
#pragma unroll
                for (int r = 0; r < R0; r++) {
                    const uint32_t h = w32[i + r * T0] >> (16 * pol);
                    const float2 v = make_float2((float)(int8_t)(h & 0xffu), (float)(int8_t)((h >> 8) & 0xffu));
                    x[r] = cmul(v, cmul(rs, sp[r * (T0 / SUB)]));
                }
                dft<R0>(x);
                if constexpr (DIF) {
                    twiddle_pow<R0>(x, args.tw, i);
                    float2* wr = buf + Gm::pad(i);
#pragma unroll
                    for (int r = 0; r < R0; r++) wr[Gm::pad(r * T0)] = x[r];
                } else {
                    float2* wr = buf + Gm::pad(i * R0);
#pragma unroll
                    for (int r = 0; r < R0; r++) wr[r] = x[r];
                }
            }
        } else {
            const uint8_t* bl = reinterpret_cast<const uint8_t*>(bytes);
            // SUB a multiple of the wave size: the step rotator index i / SUB is
            // the same in every lane of a wave, so the R0 steps b + Q r of a
            // thread (b = i / SUB, Q = T0 / SUB) are read with scalar loads from
            // the tables in L2 instead of the LDS: with one butterfly per thread
            // (SPERM) from the tables' second copy of the steps, where they are
            // contiguous, all issued before one wait; otherwise from the steps
            constexpr bool SSTEP = fstage_sstep<LG, V, CPLX>(), SPERM = fstage_sperm<LG, V, CPLX>();
            constexpr int Q = T0 / SUB;
            const float2* rtab_fa = args.rtab + ((long long)fft * p.nant + ant) * Gm::RTAB;
            const float2* rstep_g = rtab_fa + (SPERM ? Gm::NTAB : SUB);
#pragma unroll
            for (int m = 0; m < PER0; m++) {
                const int i = tid + m * NT;
                const float2 rs = rsub[i & (SUB - 1)];
                // sample n = i + r*T0: T0 is even and a multiple of SUB, so the
                // time slot is fixed per thread and the byte, step-rotator and
                // LDS write addresses are a base plus r times a constant
                const int q0 = o + i;
                const float* lt = lut + (q0 & 1);
                const uint8_t* bp = bl + ((q0 >> 1) - bbase);
                const float2* sp = rstep + i / SUB;
                const float2* spg = rstep_g + __builtin_amdgcn_readfirstlane(i / SUB) * (SPERM ? R0 : 1);
                float2 x[R0];
                [[maybe_unused]] float2 stv[R0];
                if constexpr (SPERM) {
#pragma unroll
                    for (int r = 0; r < R0; r++) stv[r] = load_const(spg + r);
                    // the loads complete here, before the LDS reads below,
                    // whose waits then count LDS operations only
                    asm volatile("" ::"s"(stv[R0 - 1].x));
                }
                if constexpr (TAN) {
                    // u[n] = L[n] * rsub[s] * rstep[k].  The fine rotator is the
                    // same for the thread's R0 samples (T0 is a multiple of SUB),
                    // so it is split as rsub[s] = f (1 + i t) (tangent form): f
                    // scales the levels, the step rotators enter the DFT's first
                    // butterfly layer, (u[a], u[a + R0/2]), with the scaled levels
                    // as the multipliers of two FMAs per component, and (1 + i t)
                    // rotates the layer's outputs with two FMAs each: six
                    // operations per sample where a complex product per rotator
                    // and the level layer took seven
                    const float2 rtf = lf_form(rs);
                    float lv[R0];
                    float2 st[R0];
#pragma unroll
                    for (int r = 0; r < R0; r++) {
                        const float l = lt[2 * bp[r * (T0 / 2)]];
                        if (TAP && args.ltap)
                            args.ltap[(((long long)fft * p.nant + ant) * 2 + pol) * F + i + r * T0] = l;
                        lv[r] = l * rtf.y;
                        st[r] = SPERM ? stv[r] : SSTEP ? load_const(spg + r * Q) : sp[r * Q];
                    }
#pragma unroll
                    for (int a = 0; a < R0 / 2; a++) {
                        const float2 sa = st[a], sb = st[a + R0 / 2];
                        const float mx = lv[a + R0 / 2] * sb.x, my = lv[a + R0 / 2] * sb.y;
                        const float la = lv[a], t = rtf.x;
                        const float2 u0 = make_float2(__builtin_fmaf(la, sa.x, mx), __builtin_fmaf(la, sa.y, my));
                        const float2 u1 = make_float2(__builtin_fmaf(la, sa.x, -mx), __builtin_fmaf(la, sa.y, -my));
                        x[a] = make_float2(__builtin_fmaf(-t, u0.y, u0.x), __builtin_fmaf(t, u0.x, u0.y));
                        x[a + R0 / 2] = make_float2(__builtin_fmaf(-t, u1.y, u1.x), __builtin_fmaf(t, u1.x, u1.y));
                    }
                } else {
                    // u[n] = L[n] * (rsub[s] * rstep[k]): the rotator is one contracted
                    // complex product, and the real level enters the DFT's first
                    // butterfly layer, (u[a], u[a + R0/2]), as the multiplier of
                    // two FMAs per component: one operation per sample fewer than
                    // scaling each rotator by its level
                    float lv[R0];
#pragma unroll
                    for (int r = 0; r < R0; r++) {
                        lv[r] = lt[2 * bp[r * (T0 / 2)]];
                        if (TAP && args.ltap)
                            args.ltap[(((long long)fft * p.nant + ant) * 2 + pol) * F + i + r * T0] = lv[r];
                        const float2 st = SPERM ? stv[r] : SSTEP ? load_const(spg + r * Q) : sp[r * Q];
                        x[r] = cmul(rs, st);
                    }
#pragma unroll
                    for (int a = 0; a < R0 / 2; a++) {
                        const float2 ua = x[a], xb = x[a + R0 / 2];
                        const float la = lv[a], lb = lv[a + R0 / 2];
                        const float2 vb = make_float2(lb * xb.x, lb * xb.y);
                        x[a] = make_float2(__builtin_fmaf(la, ua.x, vb.x), __builtin_fmaf(la, ua.y, vb.y));
                        x[a + R0 / 2] = make_float2(__builtin_fmaf(la, ua.x, -vb.x), __builtin_fmaf(la, ua.y, -vb.y));
                    }
                }
                dft<R0, false, true>(x);
                if constexpr (DIF) {
                    twiddle_pow<R0>(x, args.tw, i);
                    float2* wr = buf + Gm::pad(i);
#pragma unroll
                    for (int r = 0; r < R0; r++) wr[Gm::pad(r * T0)] = x[r];
                } else {
                    float2* wr = buf + Gm::pad(i * R0);
#pragma unroll
                    for (int r = 0; r < R0; r++) wr[r] = x[r];
                }
            }
        }
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
        __syncthreads();

        if constexpr (XL) {
            mid_pass<Gm, R1, R0, true, 1, TAN>(buf, tw1, args.tw, tid);
            // third pass (radix 16 on sub-transforms of 256), butterflies k
            // and 256 + k in the two half-waves; its outputs stay in registers
            const int up = (tid >> 5) & 1;
            const int k = (tid >> 6) * 32 + (tid & 31);
            constexpr int T2 = F / R2;
            float2 x[16];
            const float2* rd = buf + Gm::pad(k + 256 * up);
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = rd[Gm::pad(r * T2)];
            if constexpr (XLH) {
#pragma unroll
                for (int r = 1; r < 16; r++) x[r] = cmul(x[r], xw3[r]);
            } else {
                twiddle_pow<16>(x, args.tw, k * (F / (256 * 16)));
            }
            dft<16>(x);
            // x[r] is A[k + 256 r] (lanes 0-31) or B[k + 256 r] (lanes 32-63):
            // trade halves so each lane holds A and B of rho = r + 8*up
            float2 av[8], bv[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const auto sx = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[r].x), __float_as_uint(x[r + 8].x),
                                                                 false, false);
                const auto sy = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[r].y), __float_as_uint(x[r + 8].y),
                                                                 false, false);
                av[r] = make_float2(__uint_as_float(sx[0]), __uint_as_float(sy[0]));
                bv[r] = make_float2(__uint_as_float(sx[1]), __uint_as_float(sy[1]));
            }
            // last radix-2 step and fractional-delay rotation; c = k + 256*rho:
            // c % 64 = k % 64, c / 64 = k / 64 + 32 up + 4 r
            const float2 frs = frsub[k & (STRIDE - 1)];
            const float2* frb = frstep + (k >> 6) + 32 * up;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const float2 wb = cmul(bv[r], XLH ? xwl[r] : twc(xl_wk, r));
                const int c = k + 256 * (r + 8 * up);
                {
                    const float2 y = cmul(cmul(frs, frb[4 * r]), cadd(av[r], wb));
                    float si = 1.0f;
                    if constexpr (SREG) si = sreg[r * RO];
                    else if (PREC == FX_PREC_F8) si = args.scale_inv[(long long)sig * NCH + c];
                    outr.put_last(r * RO, y, si, gi);
                    if (TAP) args.tap[(((long long)fft * p.nant + ant) * 2 + pol) * NCH + c] = y;
                }
                if constexpr (CPLX) {                   // channel c + F/2 = a - W b
                    const float2 y = cmul(cmul(frs, frb[4 * r + 64]), csub(av[r], wb));
                    float si = 1.0f;
                    if constexpr (SREG) si = sreg[r * RO + 1];
                    else if (PREC == FX_PREC_F8) si = args.scale_inv[(long long)sig * NCH + c + F / 2];
                    outr.put_last(r * RO + 1, y, si, gi);
                    if (TAP) args.tap[(((long long)fft * p.nant + ant) * 2 + pol) * NCH + c + F / 2] = y;
                }
            }
        } else {
        // ---- middle passes
        if constexpr (DIF) {
            dif_pass<Gm, R1, F / R0>(buf, args.tw, tid);
            if constexpr (Gm::NP == 4) dif_pass<Gm, R2, F / (R0 * R1)>(buf, args.tw, tid);
        } else {
            mid_pass<Gm, R1, R0, true, 1, TAN>(buf, tw1, args.tw, tid);
            if constexpr (Gm::NP == 4)
                mid_pass<Gm, R2, R0 * R1, false>(buf, tw1, args.tw, tid, TWREG ? fac2 : nullptr);
        }

        // ---- last pass: radix RL, the channels (real input: bins < nchan),
        // frac-sample rotation
        // Stockham: butterfly i = tid + m*NT reads pad(i + r*TL)
        // = pad(tid) + pad(m*NT) + pad(r*TL) and its channels c = i + r*PL have
        // c % STRIDE = tid % STRIDE and c / STRIDE = tid / STRIDE + (m*NT +
        // r*PL) / STRIDE, so every address is a per-thread base plus a constant
        static_assert(NT % Gm::PM == 0 && TL % Gm::PM == 0 && NT % STRIDE == 0 && PL % STRIDE == 0, "constant offsets");
        const float2* rdb = buf + Gm::pad(tid);
        const float2 frs = frsub[tid & (STRIDE - 1)];
        const float2* frb = frstep + tid / STRIDE;
#pragma unroll
        for (int m = 0; m < PERL; m++) {
            const int i = tid + m * NT;
            float2 x[RL];
            int c0;
            if constexpr (DIF) {
                // the RL consecutive points of butterfly i; no twiddle after the last DIF pass
                static_assert(Gm::PM % RL == 0, "pad(i*RL + r) = pad(i*RL) + r");
                const float2* rd = buf + Gm::pad(i * RL);
#pragma unroll
                for (int r = 0; r < RL; r++) x[r] = rd[r];
                c0 = dif_freq<Gm>(i);
            } else {
                const float2* rd = rdb + Gm::pad(m * NT);
#pragma unroll
                for (int r = 0; r < RL; r++) x[r] = rd[Gm::pad(r * TL)];
                if constexpr (TWREG) twiddle_apply<RL>(x, facl[m]);
                else twiddle_pow<RL>(x, args.tw, i);
                c0 = i;
            }
            dft<RL, !CPLX>(x);
#pragma unroll
            for (int r = 0; r < RO; r++) {
                const int c = c0 + r * PL;
                float2 fr;
                if constexpr (DIF) fr = cmul(frsub[c & (STRIDE - 1)], frstep[c / STRIDE]);
                else fr = cmul(frs, frb[(m * NT + r * PL) / STRIDE]);
                const float2 y = cmul(fr, x[r]);
                float si = 1.0f;
                if constexpr (SREG) si = sreg[m * RO + r];
                else if (PREC == FX_PREC_F8) si = args.scale_inv[(long long)sig * NCH + c];
                outr.put_last(m * RO + r, y, si, gi);
                if (TAP) args.tap[(((long long)fft * p.nant + ant) * 2 + pol) * NCH + c] = y;
            }
        }
        }   // not XL
    }

    // ---- stores: corner-turned [chunk][chan][group in chunk][Spad] rows of G FFTs
    tid = threadIdx.x;
#pragma unroll
    for (int m = 0; m < PERL; m++) {
#pragma unroll
        for (int r = 0; r < RO; r++) {
            const int c = out_chan<Gm>(tid, m, r);
            const int ou = m * RO + r;
            const long long row = xrow0 + (long long)c * cpitch;
            if constexpr (PREC == FX_PREC_F16) {
                uint4* dst = reinterpret_cast<uint4*>(args.out) + row * (G / 4);
#pragma unroll
                for (int q = 0; q < G / 4; q++)
                    dst[q] = make_uint4(outr.v[4 * q][ou], outr.v[4 * q + 1][ou], outr.v[4 * q + 2][ou],
                                        outr.v[4 * q + 3][ou]);
            } else if constexpr (PREC == FX_PREC_F32) {
                static_assert(G == 2, "FP32 rows are (re0, re1, im0, im1)");
                float4* dst = reinterpret_cast<float4*>(args.out) + row;
                *dst = make_float4(outr.v[0][ou].x, outr.v[1][ou].x, outr.v[0][ou].y, outr.v[1][ou].y);
            } else if constexpr (F8HALF) {
                static_assert(G == 8, "two half rows of four FFTs");
                reinterpret_cast<uint2*>(args.out)[row * 2 + 1] = make_uint2(outr.v[0][ou], outr.v[1][ou]);
            } else {
                uint4* dst = reinterpret_cast<uint4*>(args.out) + row * (G / 8);
#pragma unroll
                for (int q = 0; q < G / 8; q++)
                    dst[q] = make_uint4(outr.v[4 * q][ou], outr.v[4 * q + 1][ou], outr.v[4 * q + 2][ou],
                                        outr.v[4 * q + 3][ou]);
            }
        }
    }
}

// --------------------------------------------------------------------------
// F = 32768 (16384 channels).  An FP32 transform of this length does not fit
// the LDS, so it is formed from its even- and odd-sample halves,
//   X[c] = E[c] + W^c O[c],   c < F/2 (exactly the channels kept),
// each half a 16384-point transform on the F = 16384 passes.  E stays in
// registers while O is formed.  The FFT's packed bytes (F/2 + 1 of them) are
// staged in the LDS once, with coalesced word loads, and each half reads its
// time slot from there (the byte index is contiguous in the thread index).
// A block's G FFTs share each 16-byte output row, and one FFT's outputs do
// not fit in registers beside the next FFT's work, so FFTs 0 .. G-2 park
// their encoded outputs in the block's staging area in thread order (whole
// cache lines) and FFT G-1 reads them back and writes every row whole: one
// 16-byte store per row instead of G scattered 2- to 8-byte pieces (the L2
// takes one request per piece).

static constexpr int kEoLG = 15;

// One FFT's output value of one row as the row stores it: FP32 (re, im),
// FP16 half2, FP8 two e4m3 bytes (re low).
template <int PREC> struct SlotVal;
template <> struct SlotVal<FX_PREC_F32> { using T = float2; };
template <> struct SlotVal<FX_PREC_F16> { using T = uint32_t; };
template <> struct SlotVal<FX_PREC_F8>  { using T = uint16_t; };

template <int PREC>
__device__ __forceinline__ typename SlotVal<PREC>::T encode_slot(float2 y, float si)
{
    if constexpr (PREC == FX_PREC_F32) {
        return y;
    } else if constexpr (PREC == FX_PREC_F16) {
        const __half2 h = __floats2half2_rn(y.x, y.y);
        return *reinterpret_cast<const uint32_t*>(&h);
    } else {
        const float a = fminf(fmaxf(y.x * si, -448.0f), 448.0f);
        const float b = fminf(fmaxf(y.y * si, -448.0f), 448.0f);
        return (uint16_t)((uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xffffu);
    }
}

// Row of G FFTs from the first G-1 (v[0 .. G-2]) and the last (y)
template <int PREC, int G>
__device__ __forceinline__ uint4 make_row(const typename SlotVal<PREC>::T (&v)[G - 1], typename SlotVal<PREC>::T y)
{
    if constexpr (PREC == FX_PREC_F32) {
        static_assert(G == 2, "FP32 rows: (re0, re1, im0, im1)");
        return make_uint4(__float_as_uint(v[0].x), __float_as_uint(y.x), __float_as_uint(v[0].y), __float_as_uint(y.y));
    } else if constexpr (PREC == FX_PREC_F16) {
        static_assert(G == 4, "FP16 rows: four half2");
        return make_uint4(v[0], v[1], v[2], y);
    } else {
        static_assert(G == 8, "FP8 rows: eight (re, im) byte pairs");
        return make_uint4(v[0] | (uint32_t)v[1] << 16, v[2] | (uint32_t)v[3] << 16, v[4] | (uint32_t)v[5] << 16,
                          v[6] | (uint32_t)y << 16);
    }
}

size_t fx_fstage_stage_bytes(int lg, int prec, int nant, int ngroups)
{
    if (lg != kEoLG) return 0;
    const size_t vb = prec == FX_PREC_F32 ? sizeof(float2) : prec == FX_PREC_F16 ? sizeof(uint32_t) : sizeof(uint16_t);
    const int G = fx_fstage_group_size(prec);
    return (size_t)2 * ngroups * nant * (G - 1) * ((size_t)1 << (kEoLG - 1)) * vb;
}

template <int V, int PREC, bool TAP>
__global__ __launch_bounds__((Geo<kEoLG - 1, V>::NT)) __attribute__((amdgpu_waves_per_eu(Geo<kEoLG - 1, V>::WPE)))
void fx_fstage_eo_kernel(FStageArgs args)
{
    using Gh = Geo<kEoLG - 1, V>;                               // geometry of one half
    constexpr int LG = kEoLG, F = 1 << LG, NCH = F / 2, H = F / 2;
    constexpr int NT = Gh::NT, FP = Gh::FP;
    constexpr int STRIDE = 1 << ((LG - 1) / 2), SUB = 2 * STRIDE, NSTEP = F / SUB, NFSTEP = NCH / STRIDE;
    constexpr int NTAB = SUB + NSTEP + STRIDE + NFSTEP;
    constexpr int R0 = Gh::R0, R1 = Gh::R1, R2 = Gh::R2, RL = Gh::RL;
    constexpr int T0 = H / R0, TL = H / RL, PERL = TL / NT, PL = H / RL, PER0 = T0 / NT;
    static_assert(Gh::KIND == 0 && PER0 * NT == T0, "Stockham plans, whole pass-0 butterflies per thread");
    static_assert((2 * T0) % SUB == 0, "pass-0 step rotator is a base plus r times a constant");
    static_assert(PERL * NT == PL && RL * PL == H, "last pass covers every bin of a half");
    constexpr int G = GroupSize<PREC>::G;
    constexpr int NTW1 = R0 * R1;

    __shared__ __align__(16) float2 buf[FP];
    // level of this pol at [byte][time slot] (the window keeps raw bytes)
    __shared__ __align__(16) float lut[512];
    __shared__ __align__(16) float2 tab[NTAB];
    __shared__ float2 tw1[R0 * (R1 + 1)];
    // the FFT's window: F samples span at most F/2 + 4 bytes from its aligned
    // start, F/8 + 1 words (141 + 16 KB of LDS in all)
    constexpr int NWIN = F / 8 + 1;
    __shared__ __align__(16) uint32_t win[NWIN];
    const float2* rsub = tab;
    const float2* rstep = tab + SUB;
    const float2* frsub = tab + SUB + NSTEP;
    const float2* frstep = tab + SUB + NSTEP + STRIDE;

    const FxPlan& p = args.p;
    const int nblocks = gridDim.x;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int lid = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
    int grp, ant, pol;
    fstage_work(lid, p.nant, args.ngroups, args.tile_ant, grp, ant, pol);
    const int fgo = args.fg0 + grp;
    const int sig = 2 * ant + pol;
    const int kq = fgo / args.kc;
    const long long cpitch = fx_xin_pitch(args.kc, p.Spad);
    const long long xrow0 = (long long)kq * NCH * cpitch + (long long)(fgo - kq * args.kc) * p.Spad + sig;
    int tid = threadIdx.x;

    for (int e = tid; e < R0 * (R1 + 1); e += NT) tw1[e] = tw1_entry<R0, R1, false>(args.tw, e, F);
    for (int e = tid; e < 256; e += NT) {
        const float hi = 3.3359f;
        const float lv[4] = {-hi, -1.0f, 1.0f, hi};
        lut[2 * e] = lv[(e >> (2 * pol)) & 3];
        lut[2 * e + 1] = lv[(e >> (4 + 2 * pol)) & 3];
    }
    const uint8_t* src = args.data + (long long)ant * args.data_stride;

    // The next FFT's delay entry, rotator table and window are loaded into
    // registers before this FFT's output stores are issued (loads issued
    // after them queue behind them); the delay entry (offset, valid) one FFT
    // earlier still.
    constexpr int TPT = (NTAB + NT - 1) / NT;
    constexpr int WPT = (NWIN + NT - 1) / NT;
    struct Prefetch {
        int valid, offset;
        float2 t[TPT];
        uint32_t w[WPT];
    };
    auto dload = [&](int gi) { return load_delay_ov(args.delays + (long long)(grp * G + gi) * p.nant + ant); };
    auto fetch = [&](int gi, Prefetch& f, int2 d) {
        f.offset = d.x;
        f.valid = d.y;
        if (!f.valid) return;                                  // uniform across the block
        const float2* tb = args.rtab + ((long long)(grp * G + gi) * p.nant + ant) * (NTAB + NSTEP);
        const int bb = (f.offset >> 1) & ~3;
        const int nd = ((((f.offset + F - 1) >> 1) - bb) >> 2) + 1;
        const uint32_t* s32 = reinterpret_cast<const uint32_t*>(src + bb);
#pragma unroll
        for (int t = 0; t < TPT; t++) {
            const int e = threadIdx.x + t * NT;
            f.t[t] = e < NTAB ? tb[e] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int w = 0; w < WPT; w++) {
            const int k = threadIdx.x + w * NT;
            f.w[w] = k < nd ? s32[k] : 0u;
        }
    };
    Prefetch cur;
    fetch(0, cur, dload(0));
    int2 dnext = dload(G > 1 ? 1 : 0);

#pragma unroll 1
    for (int gi = 0; gi < G; gi++) {
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid >= 0 && tid < NT);
        const int fft = grp * G + gi;
        __syncthreads();                                       // previous FFT done with the LDS
        float2 ev[PERL][RL];                                   // E[c], then the output y[c]; c = tid + m NT + r PL
        if (!cur.valid) {
            if (gi + 1 < G) {
                fetch(gi + 1, cur, dnext);
                if (gi + 2 < G) dnext = dload(gi + 2);
            }
#pragma unroll
            for (int m = 0; m < PERL; m++)
#pragma unroll
                for (int r = 0; r < RL; r++) ev[m][r] = make_float2(0.0f, 0.0f);
        } else {
            const int o = cur.offset;
            const int bb = (o >> 1) & ~3;
            const int nd = ((((o + F - 1) >> 1) - bb) >> 2) + 1;
#pragma unroll
            for (int t = 0; t < TPT; t++) {
                const int e = tid + t * NT;
                if (e < NTAB) tab[e] = cur.t[t];
            }
#pragma unroll
            for (int w = 0; w < WPT; w++) {
                const int k = tid + w * NT;
                if (k < nd) win[k] = cur.w[w];
            }
            __syncthreads();

#pragma unroll 1
            for (int h = 0; h < 2; h++) {
                asm volatile("" : "+v"(tid));
                __builtin_assume(tid >= 0 && tid < NT);
                // ---- pass 0 of half h: samples n = 2 (i + r T0) + h, one time slot
#pragma unroll
                for (int m0 = 0; m0 < PER0; m0++) {
                    const int i = tid + m0 * NT;
                    const int n0 = 2 * i + h;
                    const float2 rs = rsub[n0 & (SUB - 1)];
                    const float2* sp = rstep + n0 / SUB;
                    const float* lt = lut + ((o + h) & 1);
                    const uint8_t* bp = reinterpret_cast<const uint8_t*>(win) + (((o + h) >> 1) - bb) + i;
                    float2 x[R0];
#pragma unroll
                    for (int r = 0; r < R0; r++) {
                        const float l = lt[2 * bp[r * T0]];
                        if (TAP && args.ltap)
                            args.ltap[(((long long)fft * p.nant + ant) * 2 + pol) * F + 2 * (i + r * T0) + h] = l;
                        x[r] = cmul(make_float2(l * rs.x, l * rs.y), sp[r * (2 * T0 / SUB)]);
                    }
                    dft<R0>(x);
                    float2* wr = buf + Gh::pad(i * R0);
#pragma unroll
                    for (int r = 0; r < R0; r++) wr[r] = x[r];
                }
                __syncthreads();
                mid_pass<Gh, R1, R0, true, 2>(buf, tw1, args.tw, tid);
                if constexpr (Gh::NP == 4) mid_pass<Gh, R2, R0 * R1, false, 2>(buf, tw1, args.tw, tid);
                // ---- last pass of half h, every bin, into registers.  One
                // table entry per butterfly, all loaded ahead: W^(2i) for the
                // first half, W^i for the second (its square is W^(2i)); the
                // twiddles W^(2ir) are products of those, at most three deep.
                float2 lw[PERL];
#pragma unroll
                for (int m = 0; m < PERL; m++) lw[m] = args.tw[(2 - h) * (tid + m * NT)];
#pragma unroll
                for (int m = 0; m < PERL; m++) {
                    const int i = tid + m * NT;
                    float2 x[RL];
                    const float2* rd = buf + Gh::pad(i);
#pragma unroll
                    for (int r = 0; r < RL; r++) x[r] = rd[Gh::pad(r * TL)];
                    {
                        // the powers w^(2^b) are kept; w^r is the product of
                        // those of r's set bits (at most log2(RL) - 1 products)
                        constexpr int NB = clog2(RL);
                        float2 pw[NB];
                        pw[0] = h == 0 ? lw[m] : cmul(lw[m], lw[m]);
#pragma unroll
                        for (int b = 1; b < NB; b++) pw[b] = cmul(pw[b - 1], pw[b - 1]);
#pragma unroll
                        for (int r = 1; r < RL; r++) {
                            float2 w = pw[clog2(r)];
#pragma unroll
                            for (int b = 0; b < clog2(r); b++)
                                if ((r >> b) & 1) w = cmul(w, pw[b]);
                            x[r] = cmul(x[r], w);
                        }
                    }
                    dft<RL>(x);
                    if (h == 0) {
#pragma unroll
                        for (int r = 0; r < RL; r++) ev[m][r] = x[r];
                    } else {
                        // W^c = W^i * exp(-2 pi i r PL / F), PL / F = 1 / (2 RL): a
                        // multiple of 1/32 turn for RL <= 16, otherwise the table
                        // entry W^(r PL), the same in every lane (a scalar load)
                        const float2 wi = lw[m];
#pragma unroll
                        for (int r = 0; r < RL; r++) {
                            const int c = i + r * PL;
                            float2 ov;
                            if constexpr ((32 * PL) % F == 0) ov = twc(cmul(x[r], wi), r * (32 * PL / F));
                            else ov = cmul(cmul(x[r], wi), load_const(args.tw + r * PL));
                            const float2 X = cadd(ev[m][r], ov);
                            const float2 fr = cmul(frsub[c & (STRIDE - 1)], frstep[c / STRIDE]);
                            ev[m][r] = cmul(fr, X);
                        }
                    }
                }
                if (h == 0) __syncthreads();                   // buf is rewritten by the second half
            }
            if (gi + 1 < G) {
                fetch(gi + 1, cur, dnext);
                if (gi + 2 < G) dnext = dload(gi + 2);
            }
        }
        // ---- this FFT's slot of every output row: staged (FFTs 0 .. G-2, at
        // [FFT][m][r][thread] of the block's area) or, with the staged slots
        // read back, the whole row (FFT G-1)
        using SV = typename SlotVal<PREC>::T;
        SV* const stg = reinterpret_cast<SV*>(args.stage) + (size_t)blockIdx.x * (G - 1) * NCH;
#pragma unroll
        for (int m = 0; m < PERL; m++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                const int c = tid + m * NT + r * PL;
                float si = 1.0f;
                if (PREC == FX_PREC_F8) si = args.scale_inv[(long long)sig * NCH + c];
                const SV v = encode_slot<PREC>(ev[m][r], si);
                if (gi < G - 1) {
                    stg[((gi * PERL + m) * RL + r) * NT + tid] = v;
                } else {
                    SV pv[G - 1];
#pragma unroll
                    for (int g = 0; g < G - 1; g++) pv[g] = stg[((g * PERL + m) * RL + r) * NT + tid];
                    reinterpret_cast<uint4*>(args.out)[xrow0 + (long long)c * cpitch] = make_row<PREC, G>(pv, v);
                }
                if (TAP) args.tap[(((long long)fft * p.nant + ant) * 2 + pol) * NCH + c] = ev[m][r];
            }
    }
}

template <int V, int PREC>
static void launch_eo_prec(const FStageArgs& a, dim3 grid, bool tap, hipStream_t s)
{
    static std::atomic<int> tile{0};
    const dim3 block(Geo<kEoLG - 1, V>::NT);
    FStageArgs b = a;
    b.tile_ant = fstage_tile_ant(tile, reinterpret_cast<const void*>(&fx_fstage_eo_kernel<V, PREC, false>),
                                 (int)block.x, s);
    if (tap) fx_fstage_eo_kernel<V, PREC, true><<<grid, block, 0, s>>>(b);
    else fx_fstage_eo_kernel<V, PREC, false><<<grid, block, 0, s>>>(b);
}

template <int V>
static hipError_t launch_eo(const FStageArgs& a, int prec, hipStream_t s)
{
    const int nblocks = 2 * a.ngroups * a.p.nant;
    if (nblocks <= 0) return hipSuccess;
    if (!a.stage) return hipErrorInvalidValue;
    const dim3 grid(nblocks);
    const bool tap = a.tap != nullptr || a.ltap != nullptr;
    switch (prec) {
    case FX_PREC_F32: launch_eo_prec<V, FX_PREC_F32>(a, grid, tap, s); break;
    case FX_PREC_F16: launch_eo_prec<V, FX_PREC_F16>(a, grid, tap, s); break;
    case FX_PREC_F8: launch_eo_prec<V, FX_PREC_F8>(a, grid, tap, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Forward C2C FFT of length 32768 (all bins) from the two halves:
// X[c] = E[c] + W^c O[c], X[c + F/2] = E[c] - W^c O[c].
__global__ __launch_bounds__(Geo<kEoLG - 1>::NT) void fx_fft_c2c_eo_kernel(const float2* __restrict__ in,
                                                                         float2* __restrict__ out,
                                                                         const float2* __restrict__ tw)
{
    using Gh = Geo<kEoLG - 1>;
    constexpr int F = 1 << kEoLG, H = F / 2, NT = Gh::NT, FP = Gh::FP;
    constexpr int R0 = Gh::R0, R1 = Gh::R1, R2 = Gh::R2, RL = Gh::RL;
    constexpr int T0 = H / R0, TL = H / RL, PERL = TL / NT, PL = H / RL;
    __shared__ __align__(16) float2 buf[FP];
    __shared__ float2 tw1[R0 * (R1 + 1)];
    const float2* x_in = in + (long long)blockIdx.x * F;
    float2* x_out = out + (long long)blockIdx.x * F;
    int tid = threadIdx.x;
    for (int e = tid; e < R0 * (R1 + 1); e += NT) tw1[e] = tw1_entry<R0, R1, false>(tw, e, F);
    float2 ev[PERL][RL];
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid >= 0 && tid < NT);
        {
            float2 x[R0];
#pragma unroll
            for (int r = 0; r < R0; r++) x[r] = x_in[2 * (tid + r * T0) + h];
            dft<R0>(x);
#pragma unroll
            for (int r = 0; r < R0; r++) buf[Gh::pad(tid * R0 + r)] = x[r];
        }
        __syncthreads();
        mid_pass<Gh, R1, R0, true, 2>(buf, tw1, tw, tid);
        mid_pass<Gh, R2, R0 * R1, false, 2>(buf, tw1, tw, tid);
#pragma unroll
        for (int m = 0; m < PERL; m++) {
            const int i = tid + m * NT;
            float2 x[RL];
#pragma unroll
            for (int r = 0; r < RL; r++) x[r] = buf[Gh::pad(i + r * TL)];
            twiddle_pow<RL>(x, tw, 2 * i);
            dft<RL>(x);
            if (h == 0) {
#pragma unroll
                for (int r = 0; r < RL; r++) ev[m][r] = x[r];
            } else {
                const float2 wi = tw[i];
#pragma unroll
                for (int r = 0; r < RL; r++) {
                    const float2 ov = twc(cmul(x[r], wi), r * (32 * PL / F));
                    x_out[i + r * PL] = cadd(ev[m][r], ov);
                    x_out[i + r * PL + H] = csub(ev[m][r], ov);
                }
            }
        }
        __syncthreads();
    }
    (void)FP;
}

// --------------------------------------------------------------------------
// Generic forward C2C FFT (ippsFFTFwd_CToC_32fc semantics: unnormalised, all
// F bins), one workgroup per transform; same passes as the F-stage.

template <int LG>
__global__ __launch_bounds__(Geo<LG>::NT) void fx_fft_c2c_kernel(const float2* __restrict__ in,
                                                                 float2* __restrict__ out,
                                                                 const float2* __restrict__ tw)
{
    // one block per transform, with the F-stage's passes
    using Gm = Geo<LG>;
    constexpr int F = Gm::F, NT = Gm::NT, FP = Gm::FP;
    constexpr int R0 = Gm::R0, R1 = Gm::R1, R2 = Gm::R2, RL = Gm::RL, PL = Gm::PL;
    constexpr int T0 = F / R0, PER0 = T0 / NT, TL = F / RL, PERL = TL / NT;
    __shared__ __align__(16) float2 buf[FP];
    __shared__ float2 tw1[R0 * (R1 + 1)];
    const float2* x_in = in + (long long)blockIdx.x * F;
    float2* x_out = out + (long long)blockIdx.x * F;
    const int tid = threadIdx.x;
    for (int e = tid; e < R0 * (R1 + 1); e += NT) tw1[e] = tw1_entry<R0, R1, false>(tw, e, F);
#pragma unroll
    for (int m = 0; m < PER0; m++) {
        const int i = tid + m * NT;
        float2 x[R0];
#pragma unroll
        for (int r = 0; r < R0; r++) x[r] = x_in[i + r * T0];
        dft<R0>(x);
#pragma unroll
        for (int r = 0; r < R0; r++) buf[Gm::pad(i * R0 + r)] = x[r];
    }
    __syncthreads();
    mid_pass<Gm, R1, R0, true>(buf, tw1, tw, tid);
    if constexpr (Gm::NP == 4) mid_pass<Gm, R2, R0 * R1, false>(buf, tw1, tw, tid);
#pragma unroll
    for (int m = 0; m < PERL; m++) {
        const int i = tid + m * NT;
        float2 x[RL];
#pragma unroll
        for (int r = 0; r < RL; r++) x[r] = buf[Gm::pad(i + r * TL)];
        twiddle_pow<RL>(x, tw, i);
        dft<RL>(x);
#pragma unroll
        for (int r = 0; r < RL; r++) x_out[i + r * PL] = x[r];
    }
    (void)FP;
}

template <int LG>
static hipError_t launch_fft_lg(const float2* in, float2* out, const float2* tw, int batch, hipStream_t s)
{
    fx_fft_c2c_kernel<LG><<<batch, Geo<LG>::NT, 0, s>>>(in, out, tw);
    return hipGetLastError();
}

hipError_t launch_fft_c2c(const float2* in, float2* out, const float2* tw, int lg, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    switch (lg) {
    case 9:  return launch_fft_lg<9>(in, out, tw, batch, s);
    case 10: return launch_fft_lg<10>(in, out, tw, batch, s);
    case 11: return launch_fft_lg<11>(in, out, tw, batch, s);
    case 12: return launch_fft_lg<12>(in, out, tw, batch, s);
    case 13: return launch_fft_lg<13>(in, out, tw, batch, s);
    case 14: return launch_fft_lg<14>(in, out, tw, batch, s);
    case 15:
        fx_fft_c2c_eo_kernel<<<batch, Geo<kEoLG - 1>::NT, 0, s>>>(in, out, tw);
        return hipGetLastError();
    default: return hipErrorInvalidValue;
    }
}

bool fstage_supported(int lg) { return lg >= 9 && lg <= 15; }

template <int LG, int V, int PREC, bool CPLX>
static void launch_lgvp(const FStageArgs& a, dim3 grid, bool tap, hipStream_t s)
{
    static std::atomic<int> tile{0};
    const dim3 block(Geo<LG, V>::NT);
    FStageArgs b = a;
    b.tile_ant = fstage_tile_ant(tile, reinterpret_cast<const void*>(&fx_fstage_kernel<LG, V, PREC, false, CPLX>),
                                 (int)block.x, s);
    if (tap) fx_fstage_kernel<LG, V, PREC, true, CPLX><<<grid, block, 0, s>>>(b);
    else fx_fstage_kernel<LG, V, PREC, false, CPLX><<<grid, block, 0, s>>>(b);
}

template <int LG, int V, bool CPLX>
static hipError_t launch_lgv(const FStageArgs& a, int prec, hipStream_t s)
{
    const int nblocks = 2 * a.ngroups * a.p.nant;     // one block per polarisation
    if (nblocks <= 0) return hipSuccess;
    const dim3 grid(nblocks);
    const bool tap = a.tap != nullptr || a.ltap != nullptr;
    switch (prec) {
    case FX_PREC_F32: launch_lgvp<LG, V, FX_PREC_F32, CPLX>(a, grid, tap, s); break;
    case FX_PREC_F16: launch_lgvp<LG, V, FX_PREC_F16, CPLX>(a, grid, tap, s); break;
    case FX_PREC_F8: launch_lgvp<LG, V, FX_PREC_F8, CPLX>(a, grid, tap, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LG>
static hipError_t launch_lg(const FStageArgs& a, int prec, hipStream_t s)
{
    if (a.p.iscomplex) {
        if constexpr (nplans(LG) > 2)
            if (a.plan == 2) return launch_lgv<LG, 2, true>(a, prec, s);
        if constexpr (nplans(LG) > 1)
            if (a.plan == 1) return launch_lgv<LG, 1, true>(a, prec, s);
        return launch_lgv<LG, 0, true>(a, prec, s);
    }
    if constexpr (nplans(LG) > 3)
        if (a.plan == 3) return launch_lgv<LG, 3, false>(a, prec, s);
    if constexpr (nplans(LG) > 2)
        if (a.plan == 2) return launch_lgv<LG, 2, false>(a, prec, s);
    if constexpr (nplans(LG) > 1)
        if (a.plan == 1) return launch_lgv<LG, 1, false>(a, prec, s);
    return launch_lgv<LG, 0, false>(a, prec, s);
}

// the 32768-point halves use the Stockham plans of F = 16384 (plans 0, 1 and,
// as plan 2, the radix-32 one); complex input at F = 16384 (twice the
// channels) not the radix-32 plan
int fstage_plans(int lg, int iscomplex) { return lg == kEoLG ? 3 : (lg == 14 && iscomplex) ? 3 : nplans(lg); }

hipError_t launch_fstage(const FStageArgs& a, int prec, hipStream_t s)
{
    if (a.plan < 0 || a.plan >= fstage_plans(a.p.lg, a.p.iscomplex)) return hipErrorInvalidValue;
    switch (a.p.lg) {
    case 9:  return launch_lg<9>(a, prec, s);
    case 10: return launch_lg<10>(a, prec, s);
    case 11: return launch_lg<11>(a, prec, s);
    case 12: return launch_lg<12>(a, prec, s);
    case 13: return launch_lg<13>(a, prec, s);
    case 14: return launch_lg<14>(a, prec, s);
    case 15:
        if (a.p.iscomplex) return hipErrorInvalidValue;
        return a.plan == 2 ? launch_eo<3>(a, prec, s) : a.plan == 1 ? launch_eo<1>(a, prec, s) : launch_eo<0>(a, prec, s);
    default: return hipErrorInvalidValue;
    }
}

bool fstage_supported_complex(int lg) { return lg >= 9 && lg <= 14; }

}  // namespace fx

// fx_xengine.hip -- X-engine of the MI355X FX engine (gfx950).
//
// Per channel c the visibilities are the Hermitian matrix V = X X^H, where
// row s of X holds the spectra of signal s = 2*antenna + pol over the FFTs of
// the call.  This is the reference's processBaselinePara (kernel.c:829-854):
// V[2j+l][2k+m][c] = sum_fft chan[j][l][c] * conj(chan[k][m][c]), for every
// pair including j == k (autocorrelations, which the reference omits).
//
// V is built from 32x32 signal blocks (I, J), I <= J, on MFMA with FP32
// accumulation.  The K dimension interleaves (re, im) of consecutive FFTs:
//   Re V = sum_k A[i][k] B[j][k]       with A = B = (re, im, re, im, ...)
//   Im V = sum_k A'[i][k] B[j][k]      with A' = (im, -re, im, -re, ...)
// so each K step is two MFMAs sharing the B operand.  A and B use the same
// lane -> k assignment, so the sum is over every k whatever order the
// instruction reduces them in.
//
// A wave owns one row block I and up to NBW column blocks J0..J0+NBW-1, with
// the accumulators resident for the whole FFT range of the launch; every lane
// loads its K slice directly with one 16-byte load per block (the F-stage
// wrote the corner-turned [chan][group][signal][G] layout for this).
#include <hip/hip_runtime.h>

#include "fx_common.h"
#include "fx_kernels.h"

namespace fx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int NBW = 4;          // column blocks per wave task
constexpr int WAVES = 4;        // wave tasks per workgroup

// (re, im) fp16 pairs -> (im, -re)
__device__ __forceinline__ uint4 conj_swap_f16(uint4 x)
{
    auto f = [](uint32_t w) -> uint32_t { return ((w >> 16) | (w << 16)) ^ 0x80000000u; };
    return make_uint4(f(x.x), f(x.y), f(x.z), f(x.w));
}

// (re, im) fp8 byte pairs -> (im, -re)
__device__ __forceinline__ uint32_t conj_swap_f8w(uint32_t w)
{
    return (((w >> 8) & 0x00ff00ffu) | ((w << 8) & 0xff00ff00u)) ^ 0x80008000u;
}

__device__ __forceinline__ half8 as_half8(uint4 v)
{
    half8 h;
    __builtin_memcpy(&h, &v, sizeof(h));
    return h;
}

__device__ __forceinline__ void store_block(const XArgs& a, int c, int bp, int I, int J, int lane,
                                            const f32x16& re, const f32x16& im, const float* sc)
{
    const int n = lane & 31, kg = lane >> 5;
    float2* dst = a.vacc + ((long long)c * a.p.nbp + bp) * (FX_BLK * FX_BLK);
    float sn = 1.0f;
    if (sc) sn = sc[FX_BLK * J + n];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int m = (v >> 2) * 8 + kg * 4 + (v & 3);
        float2 val = make_float2(re[v], im[v]);
        if (sc) {
            const float s = sc[FX_BLK * I + m] * sn;
            val.x *= s;
            val.y *= s;
        }
        float2* e = dst + m * FX_BLK + n;
        if (a.accumulate) {
            const float2 o = *e;
            val.x += o.x;
            val.y += o.y;
        }
        *e = val;
    }
}

template <int PREC>
__global__ __launch_bounds__(64 * WAVES) void fx_xengine_kernel(XArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntg = (a.ntasks + a.ngroups_wg - 1) / a.ngroups_wg;
    // XCD-aware order: the task groups of one channel share an L2
    const int nblocks = gridDim.x;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int lid = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
    const int c = lid / ntg, tg = lid - c * ntg;
    const int task = tg * a.ngroups_wg + wave;
    if (task >= a.ntasks || wave >= a.ngroups_wg) return;
    const int4 t = a.tasks[task];
    const int I = t.x, J0 = t.y, nb = t.z, bp0 = t.w;
    const int row = lane & 31, kg = lane >> 5;
    const int Spad = a.p.Spad;

    f32x16 accr[NBW], acci[NBW];
#pragma unroll
    for (int b = 0; b < NBW; b++) {
        accr[b] = (f32x16){};
        acci[b] = (f32x16){};
    }

    if constexpr (PREC == FX_PREC_F16) {
        // rows of 16 bytes: (re,im) fp16 for G = 4 FFTs
        const uint4* base = reinterpret_cast<const uint4*>(a.xin) + ((long long)c * a.nfg + a.fg0) * Spad;
        const int nks = a.ngroups >> 1;          // K step = 2 groups = 8 FFTs
        const uint4* pa = base + (long long)kg * Spad + FX_BLK * I + row;
        const uint4* pb = base + (long long)kg * Spad + FX_BLK * J0 + row;
        uint4 A = pa[0];
        uint4 B[NBW];
#pragma unroll
        for (int b = 0; b < NBW; b++) B[b] = (b < nb) ? pb[FX_BLK * b] : make_uint4(0, 0, 0, 0);
        for (int ks = 0; ks < nks; ks++) {
            uint4 An = A, Bn[NBW];
#pragma unroll
            for (int b = 0; b < NBW; b++) Bn[b] = B[b];
            if (ks + 1 < nks) {
                const long long off = (long long)(ks + 1) * 2 * Spad;
                An = pa[off];
#pragma unroll
                for (int b = 0; b < NBW; b++)
                    if (b < nb) Bn[b] = pb[off + FX_BLK * b];
            }
            const half8 ha = as_half8(A);
            const half8 hc = as_half8(conj_swap_f16(A));
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                if (b < nb) {
                    const half8 hb = as_half8(B[b]);
                    accr[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ha, hb, accr[b], 0, 0, 0);
                    acci[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hc, hb, acci[b], 0, 0, 0);
                }
            }
            A = An;
#pragma unroll
            for (int b = 0; b < NBW; b++) B[b] = Bn[b];
        }
    } else if constexpr (PREC == FX_PREC_F32) {
        // rows of 16 bytes: (re0, re1, im0, im1) for G = 2 FFTs; K step = 1 FFT
        // (lanes 0-31 supply re, lanes 32-63 im)
        const float4* base = reinterpret_cast<const float4*>(a.xin) + ((long long)c * a.nfg + a.fg0) * Spad;
        const float4* pa = base + FX_BLK * I + row;
        const float4* pb = base + FX_BLK * J0 + row;
        float4 A = pa[0];
        float4 B[NBW];
#pragma unroll
        for (int b = 0; b < NBW; b++) B[b] = (b < nb) ? pb[FX_BLK * b] : make_float4(0, 0, 0, 0);
        for (int g = 0; g < a.ngroups; g++) {
            float4 An = A, Bn[NBW];
#pragma unroll
            for (int b = 0; b < NBW; b++) Bn[b] = B[b];
            if (g + 1 < a.ngroups) {
                const long long off = (long long)(g + 1) * Spad;
                An = pa[off];
#pragma unroll
                for (int b = 0; b < NBW; b++)
                    if (b < nb) Bn[b] = pb[off + FX_BLK * b];
            }
            const float av[2] = {kg ? A.z : A.x, kg ? A.w : A.y};
            const float cv[2] = {kg ? -A.x : A.z, kg ? -A.y : A.w};
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                if (b < nb) {
                    const float bv[2] = {kg ? B[b].z : B[b].x, kg ? B[b].w : B[b].y};
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        accr[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[f], bv[f], accr[b], 0, 0, 0);
                        acci[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[f], bv[f], acci[b], 0, 0, 0);
                    }
                }
            }
            A = An;
#pragma unroll
            for (int b = 0; b < NBW; b++) B[b] = Bn[b];
        }
    } else {
        // FP8 e4m3: rows of 16 bytes = 8 FFTs; K step = 4 groups = 32 FFTs,
        // each lane supplies two consecutive groups (32 bytes).
        const uint4* base = reinterpret_cast<const uint4*>(a.xin) + ((long long)c * a.nfg + a.fg0) * Spad;
        const int nks = a.ngroups >> 2;
        const uint4* pa = base + (long long)(2 * kg) * Spad + FX_BLK * I + row;
        const uint4* pb = base + (long long)(2 * kg) * Spad + FX_BLK * J0 + row;
        for (int ks = 0; ks < nks; ks++) {
            const long long off = (long long)ks * 4 * Spad;
            const uint4 a0 = pa[off], a1 = pa[off + Spad];
            i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w, (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
            i32x8 cv;
#pragma unroll
            for (int q = 0; q < 8; q++) cv[q] = (int)conj_swap_f8w((uint32_t)av[q]);
#pragma unroll
            for (int b = 0; b < NBW; b++) {
                if (b < nb) {
                    const uint4 b0 = pb[off + FX_BLK * b], b1 = pb[off + Spad + FX_BLK * b];
                    const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                                      (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                    accr[b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, accr[b], 0, 0, 0, 127, 0, 127);
                    acci[b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cv, bv, acci[b], 0, 0, 0, 127, 0, 127);
                }
            }
        }
    }

    const float* sc = (PREC == FX_PREC_F8) ? a.scale + (long long)c * Spad : nullptr;
#pragma unroll
    for (int b = 0; b < NBW; b++)
        if (b < nb) store_block(a, c, bp0 + b, I, J0 + b, lane, accr[b], acci[b], sc);
}

hipError_t launch_xengine(const XArgs& a, int prec, hipStream_t s)
{
    const int ntg = (a.ntasks + a.ngroups_wg - 1) / a.ngroups_wg;
    const int nblocks = a.p.nchan * ntg;
    if (nblocks <= 0 || a.ngroups <= 0) return hipSuccess;
    const dim3 grid(nblocks), block(64 * a.ngroups_wg);
    switch (prec) {
    case FX_PREC_F32: fx_xengine_kernel<FX_PREC_F32><<<grid, block, 0, s>>>(a); break;
    case FX_PREC_F16: fx_xengine_kernel<FX_PREC_F16><<<grid, block, 0, s>>>(a); break;
    case FX_PREC_F8:  fx_xengine_kernel<FX_PREC_F8><<<grid, block, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// counts[j*nant + k] (j <= k) += number of FFTs where both antennas are valid
// (kernel.c:850; j == k counts autocorrelation FFTs).
__global__ void fx_counts_kernel(int nant, int nffts, const FxDelay* __restrict__ d, int* __restrict__ counts)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nant * nant) return;
    const int j = t / nant, k = t - j * nant;
    if (j > k) return;
    int n = 0;
    for (int f = 0; f < nffts; f++) n += d[(long long)f * nant + j].valid & d[(long long)f * nant + k].valid;
    counts[t] += n;
}

hipError_t launch_counts(int nant, int nffts, const FxDelay* delays, int* counts, hipStream_t s)
{
    const int n = nant * nant;
    fx_counts_kernel<<<(n + 255) / 256, 256, 0, s>>>(nant, nffts, delays, counts);
    return hipGetLastError();
}

// Block accumulators -> reference layout: cross[b][2l+m][c] for j<k in
// j-major order (kernel.c:831-852) and autos[j][2l+m][c].
__global__ void fx_export_kernel(FxPlan p, const float2* __restrict__ vacc, float2* __restrict__ cross,
                                 float2* __restrict__ autos)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nbl = p.nant * (p.nant - 1) / 2;
    const long long ncross = (long long)nbl * 4 * p.nchan;
    const long long nauto = (long long)p.nant * 4 * p.nchan;
    int j, k, prod, c;
    bool isauto;
    if (t < ncross) {
        c = (int)(t % p.nchan);
        const long long bp = t / p.nchan;
        prod = (int)(bp & 3);
        int b = (int)(bp >> 2);
        j = 0;
        while (b >= p.nant - 1 - j) {
            b -= p.nant - 1 - j;
            j++;
        }
        k = j + 1 + b;
        isauto = false;
    } else if (t < ncross + nauto) {
        const long long u = t - ncross;
        c = (int)(u % p.nchan);
        const long long bp = u / p.nchan;
        prod = (int)(bp & 3);
        j = k = (int)(bp >> 2);
        isauto = true;
    } else {
        return;
    }
    const int l = prod >> 1, m = prod & 1;
    int i1 = 2 * j + l, i2 = 2 * k + m;
    bool conj = false;
    if (i1 > i2) {
        const int tmp = i1; i1 = i2; i2 = tmp;
        conj = true;
    }
    const int I = i1 / FX_BLK, J = i2 / FX_BLK;
    const int bpi = fx_bp_index(I, J, p.nblk);
    float2 v = vacc[((long long)c * p.nbp + bpi) * (FX_BLK * FX_BLK) + (i1 % FX_BLK) * FX_BLK + (i2 % FX_BLK)];
    if (conj) v.y = -v.y;
    if (isauto) {
        if (autos) autos[t - ncross] = v;
    } else if (cross) {
        cross[t] = v;
    }
}

hipError_t launch_export(const FxPlan& p, const float2* vacc, float2* cross, float2* autos, hipStream_t s)
{
    const int nbl = p.nant * (p.nant - 1) / 2;
    const long long n = ((long long)nbl + p.nant) * 4 * p.nchan;
    fx_export_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, vacc, cross, autos);
    return hipGetLastError();
}

// Fill with zeros (a kernel, so graph capture records no memset node).
__global__ void fx_zero_kernel(uint4* p, long long n16, unsigned char* tail, int ntail)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long i = t; i < n16; i += step) p[i] = make_uint4(0, 0, 0, 0);
    if (t < ntail) tail[t] = 0;
}

hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s)
{
    if (!bytes) return hipSuccess;
    const long long n16 = (long long)(bytes / 16);
    const int ntail = (int)(bytes - (size_t)n16 * 16);
    long long blocks = (n16 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    fx_zero_kernel<<<(unsigned)blocks, 256, 0, s>>>(reinterpret_cast<uint4*>(ptr), n16,
                                                     reinterpret_cast<unsigned char*>(ptr) + n16 * 16, ntail);
    return hipGetLastError();
}

}  // namespace fx

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
// The F-stage wrote the corner-turned [chunk][chan][group][signal] layout of
// 16-byte rows (G FFTs of one signal), so a row holds exactly one lane's K
// slice of one signal and a K chunk of one channel is KC*Spad contiguous
// rows; K chunks of two signal panels are staged through LDS and each wave
// keeps the accumulators of up to NBW blocks resident for the whole FFT range
// of the launch.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fx_common.h"
#include "fx_kernels.h"

namespace fx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int NBW = 4;          // 32x32 blocks per wave
constexpr int WAVES = 4;        // waves per workgroup

typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

// 16-byte global load, non-temporal when NT (the line is not kept in the
// caches for reuse)
template <bool NT>
__device__ __forceinline__ uint4 ld16(const uint4* p)
{
    if constexpr (NT) {
        const u32x4v v = __builtin_nontemporal_load(reinterpret_cast<const u32x4v*>(p));
        return make_uint4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}

// (re, im) fp16 pairs -> (im, -re)
__device__ __forceinline__ uint4 conj_swap_f16(uint4 x)
{
    auto f = [](uint32_t w) -> uint32_t { return ((w >> 16) | (w << 16)) ^ 0x80000000u; };
    return make_uint4(f(x.x), f(x.y), f(x.z), f(x.w));
}

// (re, im) fp8 byte pairs -> (im, -re): one byte permute (bytes 1 0 3 2) and
// the sign flip of each pair's new high byte
__device__ __forceinline__ uint32_t conj_swap_f8w(uint32_t w)
{
    return __builtin_amdgcn_perm(w, w, 0x02030001u) ^ 0x80008000u;
}

__device__ __forceinline__ half8 as_half8(uint4 v)
{
    half8 h;
    __builtin_memcpy(&h, &v, sizeof(h));
    return h;
}

__device__ __forceinline__ float4 as_float4(uint4 v)
{
    float4 f;
    __builtin_memcpy(&f, &v, sizeof(f));
    return f;
}

__device__ __forceinline__ void store_block(const XArgs& a, int c, int bp, int I, int J, int lane,
                                            const f32x16& re, const f32x16& im, const float* sc)
{
    const int n = lane & 31, kg = lane >> 5;
    float2* dst = a.vacc + ((long long)c * a.p.nbp + bp) * (FX_BLK * FX_BLK);
    float sn = 1.0f;
    if (sc) sn = sc[FX_BLK * J + n];
    // in four quarters: 4 old values first, their loads in flight together,
    // then the stores (more at once would cost a wave of occupancy)
#pragma unroll
    for (int h = 0; h < 16; h += 4) {
        float2 old[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int m = ((h + v) >> 2) * 8 + kg * 4 + ((h + v) & 3);
            old[v] = a.accumulate ? dst[m * FX_BLK + n] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int m = ((h + v) >> 2) * 8 + kg * 4 + ((h + v) & 3);
            float2 val = make_float2(re[h + v], im[h + v]);
            if (sc) {
                const float s = sc[FX_BLK * I + m] * sn;
                val.x *= s;
                val.y *= s;
            }
            dst[m * FX_BLK + n] = make_float2(val.x + old[v].x, val.y + old[v].y);
        }
    }
}

// The 16x16 quarter (qi, qj) of block (0, 0) of channel c from a 16x16 MFMA
// accumulator: rows 4 * (lane / 16) + v, column lane % 16 of the quarter.
__device__ __forceinline__ void store_block16(const XArgs& a, int c, int lane, const f32x4& re, const f32x4& im,
                                              int qi = 0, int qj = 0)
{
    const int n = lane & 15, m0 = 4 * (lane >> 4);
    float2* dst = a.vacc + ((long long)c * a.p.nbp + fx_bp_index(0, 0, a.p.nblk)) * (FX_BLK * FX_BLK) +
                  16 * qi * FX_BLK + 16 * qj;
    float2 old[4];
#pragma unroll
    for (int v = 0; v < 4; v++) old[v] = a.accumulate ? dst[(m0 + v) * FX_BLK + n] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int v = 0; v < 4; v++) dst[(m0 + v) * FX_BLK + n] = make_float2(re[v] + old[v].x, im[v] + old[v].y);
}

// K-chunk staging: groups of G FFTs per LDS stage, per precision.
template <int PREC> struct XChunk;
template <> struct XChunk<FX_PREC_F32> { static constexpr int KC = 4; };   // 8 FFTs
template <> struct XChunk<FX_PREC_F16> { static constexpr int KC = 4; };   // 16 FFTs
template <> struct XChunk<FX_PREC_F8>  { static constexpr int KC = 8; };   // 64 FFTs

// One workgroup = one channel x one super-tile pair (SI, SJ) of SB*32-signal
// panels.  The pair's 32x32 blocks are dealt round-robin to the WAVES waves
// (at most NBW each).  K chunks of both panels are staged through LDS by all
// threads with coalesced 16-byte loads, double-buffered in steps of NSUB
// chunks: the global loads of step k+1 are in flight while the MFMAs of step
// k run.  NSUB > 1 serves grids too small to fill the GPU (few channels and
// blocks), whose workgroups would otherwise wait one memory latency per chunk.
// NTL: the input is loaded non-temporally, for grids in which one super-tile
// pair covers every signal, so that no other workgroup reads the same rows.
// H16 (FP32, SB = 1, at most 16 signals): the one block is formed as its
// top-left 16x16 quarter on v_mfma_f32_16x16x4f32, a quarter of the MFMA
// work of a 32x32 block for the same K.  Q3 (FP32, one block of 17 to 32
// signals): its quarters (0, 0), (0, 1) and (1, 1) on v_mfma_f32_16x16x4f32,
// three quarters of the work; quarter (1, 0) holds the conjugates of (0, 1),
// and no visibility is read from it (the export reads i <= j).  H8 (FP32, at
// most 8 signals): a 16x16 MFMA over two groups at once, rows 0-7 of its
// operand from one group and rows 8-15 from the next, so its diagonal 8x8
// quarters are the two groups' products (summed at the end) and the others
// products across FFTs, never stored: half the MFMAs and staged rows of H16.
// NTL also means one panel pair per channel (every signal in one diagonal
// super-tile): one panel is staged, each thread issues one branch-free load
// per chunk, and the loads run two chunks ahead in two register sets (the
// waits before the LDS stores then cover only the older chunk; a conditional
// load would make the compiler wait for every load in flight).
template <int PREC, int SB, int NSUB, int NW = WAVES, bool NTL = false, bool H16 = false, bool Q3 = false,
          bool H8 = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 4 : 2))) void fx_xengine_kernel(XArgs a)
{
    constexpr int KC = XChunk<PREC>::KC;
    constexpr int PW = SB * FX_BLK;                      // signals per panel
    // rows staged per group and panel, and panels staged: the 16x16 kernels
    // stage only the rows they read (16 for H16, 8 for H8, 32 for Q3) of their
    // one (diagonal) panel; NTL the one (diagonal) panel of its grid
    constexpr bool QB = H16 || Q3 || H8;                 // 16x16 quarters
    constexpr int NQ = Q3 ? 3 : 1;
    constexpr int PWL = H16 ? FX_BLK / 2 : H8 ? FX_BLK / 4 : PW;
    constexpr int NPL = (QB || NTL) ? 1 : 2;
    constexpr int PF = NTL ? 2 : 1;                      // chunks in flight per thread
    static_assert(PF == 1 || NSUB == 1, "two chunks in flight: one chunk per step");
    constexpr int NELEM = NPL * PWL * KC;                // 16-byte rows per chunk
    constexpr int NLD = (NSUB * NELEM + 64 * NW - 1) / (64 * NW);
    constexpr bool PART = (NSUB * NELEM) % (64 * NW) != 0;   // threads past the step's rows load nothing
    // [stage][chunk][panel][group][signal], the panels and rows staged; the
    // K-split kernels' partial-sum reduction reuses it
    // (direct-to-LDS kernels keep their second stage in an array of its own,
    // so that the compiler sees that loads into one stage do not alias reads
    // of the other)
    constexpr bool DMA = NW == 8 && SB == 4 && !NTL;
    __shared__ __align__(16) uint4 lds[DMA ? 1 : 2][NSUB][NPL][KC][PWL];
    __shared__ __align__(16) uint4 lds_b[DMA ? NSUB : 1][DMA ? NPL : 1][DMA ? KC : 1][DMA ? PWL : 1];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, tid = threadIdx.x;
    // the launch has 64*NW threads: the panel of a load slot is then a
    // compile-time function of the slot wherever whole waves share it
    __builtin_assume(tid >= 0 && tid < 64 * NW);
    const int Spad = a.p.Spad, nblk = a.p.nblk;
    const int npan = (nblk + SB - 1) / SB;
    const int npairs = npan * (npan + 1) / 2;
    // XCD-aware order: the pairs of one channel share an L2
    const int nblocks = gridDim.x;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int lid = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
    const int c = lid / npairs;
    int pr = lid - c * npairs, SI = 0;
    while (pr >= npan - SI) {
        pr -= npan - SI;
        SI++;
    }
    const int SJ = SI + pr;
    const bool diag = SI == SJ;
    if constexpr (NTL) __builtin_assume(diag);            // one panel pair per channel

    // blocks of this wave.  SB = 4 with eight waves (FP16, FP8; SUBT8): a 1x2
    // sub-tile per wave (row wave/2, columns 2*(wave%2) + j in slot j), so a
    // K step reads one A and two B fragments for two block products and
    // sixteen waves per CU hide the loads (measured against four waves with
    // 2x2 sub-tiles: the 256-antenna FP8 X-engine 2.95 -> 2.67 ms per 16
    // subints).  SB = 4 with four waves (FP32; SUBT): a 2x2 sub-tile per wave
    // (rows 2*(wave/2) + i, columns 2*(wave%2) + j in slot 2i + j).  On a
    // diagonal pair the blocks below the diagonal are idle.  SB = 2: blocks
    // dealt round-robin.  SB = 1 (KSPLIT): the pair is one block; every wave
    // accumulates it over a quarter of the K steps (K step k to wave k % NW)
    // and the four partial sums are added through the LDS at the end, instead
    // of one wave computing the block while three idle.
    constexpr bool SUBT = SB == 4 && NW == 4;
    constexpr bool SUBT8 = SB == 4 && NW == 8;
    constexpr bool KSPLIT = SB == 1;
    static_assert(!SUBT || (NW == 4 && NBW == 4), "2x2 sub-tiles of a 4x4 pair");
    int bI[NBW], bJ[NBW];
    unsigned okm = 0;                                     // bit t: slot t holds a block
    if constexpr (KSPLIT) {
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            bI[t] = 0;
            bJ[t] = 0;
        }
        okm = (SI < nblk && SJ < nblk) ? 1u : 0u;
    } else if constexpr (SUBT8) {
        // off the diagonal: row wave/2, columns 2*(wave%2) + j.  On a diagonal
        // pair the ten blocks I <= J as four 1x2 and two 1x1 sub-tiles, placed
        // so that the SIMDs (wave % 4) carry 3, 3, 2 and 2 blocks
        // (waves 0-7: rows 0 0 1 2 1 3 - -, first columns 0 2 2 2 1 3 - -,
        // widths 2 2 2 2 1 1 0 0)
        const int sh = 4 * wave;
        const int ri = diag ? (0x00312100 >> sh) & 15 : wave >> 1;
        const int c0 = diag ? (0x00312220 >> sh) & 15 : 2 * (wave & 1);
        const int nw = diag ? (0x00112222 >> sh) & 15 : 2;
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            const int bi = ri, bj = c0 + (t & 1);
            const int I = SI * SB + bi, J = SJ * SB + bj;
            const bool ok = t < nw && I < nblk && J < nblk;
            bI[t] = bi;
            bJ[t] = bj;
            okm |= (unsigned)ok << t;
        }
    } else if constexpr (SUBT) {
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            const int bi = 2 * (wave >> 1) + (t >> 1), bj = 2 * (wave & 1) + (t & 1);
            const int I = SI * SB + bi, J = SJ * SB + bj;
            const bool ok = (!diag || bi <= bj) && I < nblk && J < nblk;
            bI[t] = bi;
            bJ[t] = bj;
            okm |= (unsigned)ok << t;
        }
    } else {
        const int nq = diag ? SB * (SB + 1) / 2 : SB * SB;
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            const int q = wave + t * NW;
            int bi = 0, bj = 0;
            if (diag) {
                int r = q;
                while (bi < SB && r >= SB - bi) {
                    r -= SB - bi;
                    bi++;
                }
                bj = bi + r;
            } else {
                bi = q / SB;
                bj = q % SB;
            }
            const int I = SI * SB + bi, J = SJ * SB + bj;
            const bool ok = q < nq && I < nblk && J < nblk;
            bI[t] = ok ? bi : 0;
            bJ[t] = ok ? bj : 0;
            okm |= (unsigned)ok << t;
        }
    }

    // chunk q of channel c: KC*Spad contiguous rows at ((q*nchan + c)*KC)*Spad
    const long long cpitch = fx_xin_pitch(KC, Spad), cstride = (long long)a.p.nchan * cpitch;
    const uint4* gbase = reinterpret_cast<const uint4*>(a.xin) + ((long long)(a.fg0 / KC) * a.p.nchan + c) * cpitch;
    const int nchunk = a.ngroups / KC;
    const int nstep = (nchunk + NSUB - 1) / NSUB;
    // per-thread loads of step sp: element e = (chunk in step, panel, group,
    // signal), signal fastest.  No divergent branch or zero fill per load: a
    // signal past Spad = 32*nblk reads the last row instead (it only feeds
    // blocks beyond nblk, which okm excludes), and with one chunk per step a
    // diagonal workgroup skips its second panel (a workgroup- or wave-uniform
    // branch; that LDS is not read).
    static_assert(QB || NELEM % (64 * NW) == 0, "whole loads per thread");
    uint4 st[PF][NLD];
#pragma unroll
    for (int q = 0; q < PF; q++)
#pragma unroll
        for (int l = 0; l < NLD; l++) st[q][l] = make_uint4(0, 0, 0, 0);
    auto gload = [&](int sp, uint4 (&r)[NLD]) {
#pragma unroll
        for (int l = 0; l < NLD; l++) {
            const int e = tid + l * 64 * NW;
            if (PART && e >= NSUB * NELEM) continue;
            const int sub = e / NELEM, e2 = e - sub * NELEM;
            const int pan = NPL == 1 ? 0 : e2 / (KC * PWL), rem = e2 - pan * (KC * PWL);
            const int g = rem / PWL, sgl = rem - g * PWL;
            const int sig = (pan ? SJ : SI) * PW + sgl;
            // the last step's chunks past nchunk re-read the last chunk; with
            // several chunks per step every load is issued, so that they go
            // out back to back (a branch between them serialises them)
            const int chunk = sp * NSUB + sub;
            const int chc = chunk < nchunk ? chunk : nchunk - 1;
            // a wave-uniform chunk base and a 32-bit per-thread byte offset
            // (the scalar-base load form: no per-load address in registers)
            const char* cb = reinterpret_cast<const char*>(gbase + chc * cstride);
            const unsigned boff = (unsigned)(g * Spad + (sig < Spad ? sig : Spad - 1)) * 16u;
            if (NPL == 1 || NSUB > 1 || !(diag && pan))
                r[l] = ld16<NTL>(reinterpret_cast<const uint4*>(cb + boff));
        }
    };
    auto sstore = [&](int stage, const uint4 (&r)[NLD]) {
#pragma unroll
        for (int l = 0; l < NLD; l++) {
            const int e = tid + l * 64 * NW;
            if (PART && e >= NSUB * NELEM) continue;
            const int sub = e / NELEM, e2 = e - sub * NELEM;
            const int pan = NPL == 1 ? 0 : e2 / (KC * PWL), rem = e2 - pan * (KC * PWL);
            const int g = rem / PWL, sgl = rem - g * PWL;
            lds[stage][sub][pan][g][sgl] = r[l];
        }
    };

    // Eight waves with one chunk in flight: each chunk goes from global memory
    // straight into its LDS stage (global_load_lds_dwordx4: lane i of a wave
    // writes the 16 bytes at the wave's LDS base + 16 i, and the elements e of
    // a wave are 64 consecutive rows of the stage), so no staging registers and
    // no LDS stores; a step waits for them before its closing barrier.
    static_assert(!DMA || (SUBT8 && PF == 1), "direct-to-LDS staging: eight waves, one chunk in flight");
    typedef __attribute__((address_space(3))) void lds_void;
    auto dma = [&](int sp, auto& stage) {
#pragma unroll
        for (int l = 0; l < NLD; l++) {
            const int e = tid + l * 64 * NW;
            if (PART && e >= NSUB * NELEM) continue;
            const int sub = e / NELEM, e2 = e - sub * NELEM;
            const int pan = NPL == 1 ? 0 : e2 / (KC * PWL), rem = e2 - pan * (KC * PWL);
            const int g = rem / PWL, sgl = rem - g * PWL;
            const int sig = (pan ? SJ : SI) * PW + sgl;
            const int chunk = sp * NSUB + sub;
            const int chc = chunk < nchunk ? chunk : nchunk - 1;
            const char* cb = reinterpret_cast<const char*>(gbase + chc * cstride);
            const unsigned boff = (unsigned)(g * Spad + (sig < Spad ? sig : Spad - 1)) * 16u;
            if (NPL == 1 || NSUB > 1 || !(diag && pan)) {
                uint4* wb = &stage[0][0][0][0] + (e - lane);
                __builtin_amdgcn_global_load_lds(cb + boff, (lds_void*)wb, 16, 0, NTL ? 2 : 0);
            }
        }
    };

    static_assert(!QB || (PREC == FX_PREC_F32 && SB == 1 && (int)H16 + (int)Q3 + (int)H8 == 1),
                  "16x16 blocks: FP32, one block per pair");
    f32x16 accr[NBW], acci[NBW];
#pragma unroll
    for (int t = 0; t < NBW; t++) {
        accr[t] = (f32x16){};
        acci[t] = (f32x16){};
    }
    f32x4 acr16[NQ], aci16[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        acr16[q] = (f32x4){};
        aci16[q] = (f32x4){};
    }
    const int row = lane & 31, kg = lane >> 5;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int pb = (NPL == 1 || diag) ? 0 : 1;            // panel of the B operand

    if (nstep > 0) {
        if constexpr (DMA) {
            dma(0, lds[0]);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        } else {
            gload(0, st[0]);
            sstore(0, st[0]);
        }
    }
    // (launches have at least one chunk; past the last one, the last again)
    if constexpr (PF == 2) gload(1, st[1]);
    __syncthreads();
    // KSPLIT: K step k of the launch belongs to wave k % NW
    auto mine = [&](int k) { return !KSPLIT || (k & (NW - 1)) == wave; };
    // the wave index as a wave-uniform value, for LDS addresses formed from it
    const int wvu = __builtin_amdgcn_readfirstlane(wave);
    // the MFMAs of the chunks of step sp from LDS stage stg
    auto compute = [&](int sp, auto& L) {
#pragma unroll
        for (int sub = 0; sub < NSUB; sub++) {
        if (sp * NSUB + sub >= nchunk) break;             // last step of a chunk count not a multiple of NSUB
        if constexpr (PREC == FX_PREC_F16 && SUBT8) {
            // every fragment of the chunk is read before its MFMAs, so the
            // LDS reads overlap (a slot without a block reads the first
            // column again, inside the panel)
            if (!(okm & 3u)) continue;
            const int r0 = bI[0] * FX_BLK + row;
            int cj[2];
#pragma unroll
            for (int j = 0; j < 2; j++) cj[j] = ((okm >> j) & 1u ? bJ[j] : bJ[0]) * FX_BLK + row;
            uint4 A[KC / 2], B[KC / 2][2];
#pragma unroll
            for (int ks = 0; ks < KC / 2; ks++) {
                A[ks] = L[sub][0][2 * ks + kg][r0];
#pragma unroll
                for (int j = 0; j < 2; j++) B[ks][j] = L[sub][pb][2 * ks + kg][cj[j]];
            }
#pragma unroll
            for (int ks = 0; ks < KC / 2; ks++) {
                const half8 ha = as_half8(A[ks]), hc = as_half8(conj_swap_f16(A[ks]));
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if ((okm >> j) & 1u) {
                        const half8 hb = as_half8(B[ks][j]);
                        accr[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ha, hb, accr[j], 0, 0, 0);
                        acci[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hc, hb, acci[j], 0, 0, 0);
                    }
                }
            }
        } else if constexpr (PREC == FX_PREC_F16) {
#pragma unroll
            for (int ks = 0; ks < KC / 2; ks++) {
                if (!mine((sp * NSUB + sub) * (KC / 2) + ks)) continue;
#pragma unroll
                for (int t = 0; t < NBW; t++) {
                    if ((okm >> t) & 1u) {
                        const uint4 A = L[sub][0][2 * ks + kg][bI[t] * FX_BLK + row];
                        const uint4 B = L[sub][pb][2 * ks + kg][bJ[t] * FX_BLK + row];
                        const half8 hb = as_half8(B);
                        accr[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_half8(A), hb, accr[t], 0, 0, 0);
                        acci[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_half8(conj_swap_f16(A)), hb, acci[t], 0,
                                                                         0, 0);
                    }
                }
            }
        } else if constexpr (PREC == FX_PREC_F32 && H16) {
            // one K step per row (2 FFTs), and with KC == NW the K step of
            // group g of every chunk is wave g's: lane l reads element l / 16
            // of row l % 16 of its wave's group, in the order (re0, re1, im0,
            // im1), and the element of A' = (im, -re), in the order (im0, im1,
            // -re0, -re1).  The one block is diagonal: B = A.
            static_assert(KC == NW, "one group of each chunk per wave");
            if (okm & 1u) {
                const float* rp = reinterpret_cast<const float*>(&L[sub][0][wvu][r16]);
                const float av = rp[k4];
                const float cw = rp[(k4 + 2) & 3];
                const float cv = __int_as_float(__float_as_int(cw) ^ (k4 >= 2 ? (int)0x80000000 : 0));
                acr16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, av, acr16[0], 0, 0, 0);
                aci16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv, av, aci16[0], 0, 0, 0);
            }
        } else if constexpr (PREC == FX_PREC_F32 && H8) {
            // item q of chunk sub: groups 2q (operand rows 0-7) and 2q + 1
            // (rows 8-15); the items of a step are dealt to the waves in turn
            static_assert(KC == NW && KC % 2 == 0, "pairs of groups");
            if (okm & 1u) {
#pragma unroll
                for (int q = 0; q < KC / 2; q++) {
                    if (((sub * (KC / 2) + q) & (NW - 1)) != wvu) continue;
                    const float* rp = reinterpret_cast<const float*>(&L[sub][0][2 * q + (r16 >> 3)][r16 & 7]);
                    const float av = rp[k4];
                    const float cw = rp[(k4 + 2) & 3];
                    const float cv = __int_as_float(__float_as_int(cw) ^ (k4 >= 2 ? (int)0x80000000 : 0));
                    acr16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, av, acr16[0], 0, 0, 0);
                    aci16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv, av, aci16[0], 0, 0, 0);
                }
            }
        } else if constexpr (PREC == FX_PREC_F32 && Q3) {
            // as H16, for rows r16 (quarter row 0) and 16 + r16 (quarter row
            // 1): quarter (qi, qj) is rows qi against rows qj
            static_assert(KC == NW, "one group of each chunk per wave");
            if (okm & 1u) {
                const float* rp0 = reinterpret_cast<const float*>(&L[sub][0][wvu][r16]);
                const float* rp1 = reinterpret_cast<const float*>(&L[sub][0][wvu][16 + r16]);
                const int sg = k4 >= 2 ? (int)0x80000000 : 0;
                const float a0 = rp0[k4], a1 = rp1[k4];
                const float c0 = __int_as_float(__float_as_int(rp0[(k4 + 2) & 3]) ^ sg);
                const float c1 = __int_as_float(__float_as_int(rp1[(k4 + 2) & 3]) ^ sg);
                acr16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, a0, acr16[0], 0, 0, 0);
                aci16[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0, a0, aci16[0], 0, 0, 0);
                acr16[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, a1, acr16[1], 0, 0, 0);
                aci16[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0, a1, aci16[1], 0, 0, 0);
                acr16[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, a1, acr16[2], 0, 0, 0);
                aci16[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(c1, a1, aci16[2], 0, 0, 0);
            }
        } else if constexpr (PREC == FX_PREC_F32) {
            // rows (re0, re1, im0, im1); K step = 1 FFT, lanes 0-31 re, 32-63 im
#pragma unroll
            for (int g = 0; g < KC; g++) {
                if (!mine((sp * NSUB + sub) * KC + g)) continue;
#pragma unroll
                for (int t = 0; t < NBW; t++) {
                    if ((okm >> t) & 1u) {
                        const float4 A = as_float4(L[sub][0][g][bI[t] * FX_BLK + row]);
                        const float4 B = as_float4(L[sub][pb][g][bJ[t] * FX_BLK + row]);
                        const float av[2] = {kg ? A.z : A.x, kg ? A.w : A.y};
                        const float cv[2] = {kg ? -A.x : A.z, kg ? -A.y : A.w};
                        const float bv[2] = {kg ? B.z : B.x, kg ? B.w : B.y};
#pragma unroll
                        for (int f = 0; f < 2; f++) {
                            accr[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[f], bv[f], accr[t], 0, 0, 0);
                            acci[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(cv[f], bv[f], acci[t], 0, 0, 0);
                        }
                    }
                }
            }
        } else if constexpr (SUBT8 && PREC == FX_PREC_F8) {
            const int r0 = bI[0] * FX_BLK + row, c0 = bJ[0] * FX_BLK + row;
#pragma unroll
            for (int ks = 0; ks < KC / 4; ks++) {
                if (!(okm & 3u)) break;
                const uint4 a0 = L[sub][0][4 * ks + 2 * kg][r0];
                const uint4 a1 = L[sub][0][4 * ks + 2 * kg + 1][r0];
                const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                                  (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                i32x8 cv;
#pragma unroll
                for (int q = 0; q < 8; q++) cv[q] = (int)conj_swap_f8w((uint32_t)av[q]);
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if ((okm >> j) & 1u) {
                        const uint4 b0 = L[sub][pb][4 * ks + 2 * kg][c0 + j * FX_BLK];
                        const uint4 b1 = L[sub][pb][4 * ks + 2 * kg + 1][c0 + j * FX_BLK];
                        const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                                          (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                        accr[j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, accr[j], 0, 0, 0, 127, 0, 127);
                        acci[j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cv, bv, acci[j], 0, 0, 0, 127, 0, 127);
                    }
                }
            }
        } else {
            // FP8 e4m3: K step = 4 groups (32 FFTs); lane kg supplies groups 2kg, 2kg+1
#pragma unroll
            for (int ks = 0; ks < KC / 4; ks++) {
                if (!mine((sp * NSUB + sub) * (KC / 4) + ks)) continue;
#pragma unroll
                for (int t = 0; t < NBW; t++) {
                    if ((okm >> t) & 1u) {
                        const uint4 a0 = L[sub][0][4 * ks + 2 * kg][bI[t] * FX_BLK + row];
                        const uint4 a1 = L[sub][0][4 * ks + 2 * kg + 1][bI[t] * FX_BLK + row];
                        const uint4 b0 = L[sub][pb][4 * ks + 2 * kg][bJ[t] * FX_BLK + row];
                        const uint4 b1 = L[sub][pb][4 * ks + 2 * kg + 1][bJ[t] * FX_BLK + row];
                        const i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                                          (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
                        i32x8 cv;
#pragma unroll
                        for (int q = 0; q < 8; q++) cv[q] = (int)conj_swap_f8w((uint32_t)av[q]);
                        const i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                                          (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
                        accr[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, accr[t], 0, 0, 0, 127, 0, 127);
                        acci[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(cv, bv, acci[t], 0, 0, 0, 127, 0, 127);
                    }
                }
            }
        }
        }   // sub
    };
    if constexpr (DMA) {
        // stage 0 is lds[0], stage 1 lds_b: two steps per trip
        for (int sp = 0; sp < nstep; sp += 2) {
            if (sp + 1 < nstep) dma(sp + 1, lds_b);
            compute(sp, lds[0]);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if (sp + 1 >= nstep) break;
            if (sp + 2 < nstep) dma(sp + 2, lds[0]);
            compute(sp + 1, lds_b);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
    } else if constexpr (PF == 1) {
        for (int sp = 0; sp < nstep; sp++) {
            const int stg = sp & 1;
            if (sp + 1 < nstep) gload(sp + 1, st[0]);
            compute(sp, lds[stg]);
            if (sp + 1 < nstep) sstore(stg ^ 1, st[0]);
            __syncthreads();
        }
    } else {
        // chunk k travels in register set k & 1 and is staged in LDS stage
        // k & 1: step sp issues chunk sp + 2 into the set chunk sp left,
        // computes chunk sp, then stages chunk sp + 1.  Loads and stores run
        // unconditionally (past the last chunk they repeat it, and compute
        // skips steps past it): with every path issuing the same loads, the
        // compiler's waits before the stores cover only the older chunk.
        auto step = [&](int sp, auto par) {
            constexpr int P = decltype(par)::value;
            gload(sp + 2, st[P]);
            compute(sp, lds[P]);
            sstore(P ^ 1, st[P ^ 1]);
            __syncthreads();
        };
        for (int sp = 0; sp < nstep; sp += 2) {
            step(sp, std::integral_constant<int, 0>{});
            step(sp + 1, std::integral_constant<int, 1>{});
        }
    }

    if constexpr (KSPLIT) {
        // wave 0 adds the other waves' partial sums, one wave at a time
        // through the staging LDS (free after the loop's last barrier; it
        // holds at least one wave's 2 x 32 x 32 floats, or 2 x 16 x 16 for a
        // 16x16 quarter); the order is fixed,
        // so the result is the same on every run
        static_assert(sizeof(lds) >= 2 * (QB ? 4 : 16) * 64 * sizeof(float), "staging LDS holds one wave's partial sums");
        float* red = reinterpret_cast<float*>(&lds[0][0][0][0][0]);
#pragma unroll
        for (int u = 0; u < (QB ? NQ : 1) * (NW - 1); u++) {
            const int w = 1 + u % (NW - 1), q = QB ? u / (NW - 1) : 0;   // quarter q of wave w
            if (wave == w) {
                if constexpr (QB) {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        red[v * 64 + lane] = acr16[q][v];
                        red[(4 + v) * 64 + lane] = aci16[q][v];
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        red[v * 64 + lane] = accr[0][v];
                        red[(16 + v) * 64 + lane] = acci[0][v];
                    }
                }
            }
            __syncthreads();
            if (wave == 0) {
                if constexpr (QB) {
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        acr16[q][v] += red[v * 64 + lane];
                        aci16[q][v] += red[(4 + v) * 64 + lane];
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        accr[0][v] += red[v * 64 + lane];
                        acci[0][v] += red[(16 + v) * 64 + lane];
                    }
                }
            }
            __syncthreads();
        }
        if (wave != 0) okm = 0;
        if constexpr (H8) {
            // V[m][n] = D[m][n] + D[m + 8][n + 8] for m, n < 8: lane l holds
            // D[4 (l / 16) + v][l % 16], so D[m + 8][n + 8] is in lane l + 40;
            // every other entry is set to zero (added to the accumulators, it
            // leaves them as they are).  Wave 0 alone, through the staging LDS.
            if (wave == 0) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    red[v * 64 + lane] = acr16[0][v];
                    red[(4 + v) * 64 + lane] = aci16[0][v];
                }
                const bool keep = (lane >> 4) < 2 && (lane & 15) < 8;
                const int src = (lane + 40) & 63;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    acr16[0][v] = keep ? acr16[0][v] + red[v * 64 + src] : 0.0f;
                    aci16[0][v] = keep ? aci16[0][v] + red[(4 + v) * 64 + src] : 0.0f;
                }
            }
        }
    }
    if constexpr (H16 || H8) {
        if (okm & 1u) store_block16(a, c, lane, acr16[0], aci16[0]);
        return;
    }
    if constexpr (Q3) {
        if (okm & 1u) {
            store_block16(a, c, lane, acr16[0], aci16[0], 0, 0);
            store_block16(a, c, lane, acr16[1], aci16[1], 0, 1);
            store_block16(a, c, lane, acr16[2], aci16[2], 1, 1);
        }
        return;
    }
    const float* sc = (PREC == FX_PREC_F8) ? a.scale + (long long)c * Spad : nullptr;
#pragma unroll
    for (int t = 0; t < NBW; t++) {
        if ((okm >> t) & 1u) {
            const int I = SI * SB + bI[t], J = SJ * SB + bJ[t];
            store_block(a, c, fx_bp_index(I, J, nblk), I, J, lane, accr[t], acci[t], sc);
        }
    }
}

// Plain FP32 X-engine on the vector ALUs (no MFMA): the same workgroup
// decomposition, LDS staging and accumulator layout as the MFMA kernels, with
// each lane accumulating one column n and 16 rows m of a 32x32 block by
// scalar complex multiply-adds (the structure of the reference's CCAH3,
// gkernel.cu:298-365, register-tiled).
template <int SB>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2))) void fx_xengine_vec_kernel(XArgs a)
{
    constexpr int KC = XChunk<FX_PREC_F32>::KC;
    constexpr int PW = SB * FX_BLK;
    constexpr int NELEM = 2 * PW * KC;
    constexpr int NLD = (NELEM + 64 * WAVES - 1) / (64 * WAVES);
    constexpr int NBV = 3;                                // blocks per wave (10 blocks / 4 waves)
    __shared__ __align__(16) float4 lds[2][2][KC][PW];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, tid = threadIdx.x;
    const int Spad = a.p.Spad, nblk = a.p.nblk;
    const int npan = (nblk + SB - 1) / SB;
    const int npairs = npan * (npan + 1) / 2;
    const int nblocks = gridDim.x;
    const int q8 = nblocks >> 3, r8 = nblocks & 7;
    const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
    const int lid = xcd * q8 + (xcd < r8 ? xcd : r8) + loc;
    const int c = lid / npairs;
    int pr = lid - c * npairs, SI = 0;
    while (pr >= npan - SI) {
        pr -= npan - SI;
        SI++;
    }
    const int SJ = SI + pr;
    const bool diag = SI == SJ;
    const int nq = diag ? SB * (SB + 1) / 2 : SB * SB;
    const int pb = diag ? 0 : 1;
    const long long cpitch = fx_xin_pitch(KC, Spad), cstride = (long long)a.p.nchan * cpitch;
    const float4* gbase = reinterpret_cast<const float4*>(a.xin) + ((long long)(a.fg0 / KC) * a.p.nchan + c) * cpitch;
    const int nchunk = a.ngroups / KC;
    const int n = lane & 31, kg = lane >> 5;

    for (int q0 = 0; q0 < nq; q0 += WAVES * NBV) {
        // this wave's blocks
        int bI[NBV], bJ[NBV];
        unsigned okm = 0;
#pragma unroll
        for (int t = 0; t < NBV; t++) {
            const int q = q0 + wave + t * WAVES;
            int bi = 0, bj = 0;
            if (diag) {
                int r = q;
                while (bi < SB && r >= SB - bi) {
                    r -= SB - bi;
                    bi++;
                }
                bj = bi + r;
            } else {
                bi = q / SB;
                bj = q % SB;
            }
            const bool ok = q < nq && SI * SB + bi < nblk && SJ * SB + bj < nblk;
            bI[t] = ok ? bi : 0;
            bJ[t] = ok ? bj : 0;
            okm |= (unsigned)ok << t;
        }
        float accr[NBV][16], acci[NBV][16];
#pragma unroll
        for (int t = 0; t < NBV; t++)
#pragma unroll
            for (int v = 0; v < 16; v++) accr[t][v] = acci[t][v] = 0.0f;
        float4 st[NLD];
        auto gload = [&](int chunk) {
#pragma unroll
            for (int l = 0; l < NLD; l++) {
                const int e = tid + l * 64 * WAVES;
                const int pan = e / (KC * PW), rem = e - pan * (KC * PW);
                const int g = rem / PW, sgl = rem - g * PW;
                const int sig = (pan ? SJ : SI) * PW + sgl;
                st[l] = make_float4(0, 0, 0, 0);
                if (e < NELEM && sig < Spad && !(diag && pan)) st[l] = gbase[chunk * cstride + g * Spad + sig];
            }
        };
        auto sstore = [&](int stage) {
#pragma unroll
            for (int l = 0; l < NLD; l++) {
                const int e = tid + l * 64 * WAVES;
                if (e < NELEM) {
                    const int pan = e / (KC * PW), rem = e - pan * (KC * PW);
                    const int g = rem / PW, sgl = rem - g * PW;
                    lds[stage][pan][g][sgl] = st[l];
                }
            }
        };
        __syncthreads();
        if (nchunk > 0) {
            gload(0);
            sstore(0);
        }
        __syncthreads();
        for (int ch = 0; ch < nchunk; ch++) {
            const int stg = ch & 1;
            if (ch + 1 < nchunk) gload(ch + 1);
#pragma unroll
            for (int g = 0; g < KC; g++) {
#pragma unroll
                for (int t = 0; t < NBV; t++) {
                    if ((okm >> t) & 1u) {
                        const float4 B = lds[stg][pb][g][bJ[t] * FX_BLK + n];
#pragma unroll
                        for (int v = 0; v < 16; v++) {
                            const int m = (v >> 2) * 8 + kg * 4 + (v & 3);
                            const float4 A = lds[stg][0][g][bI[t] * FX_BLK + m];
                            // A * conj(B) for the two FFTs of the row
                            accr[t][v] += A.x * B.x + A.z * B.z;
                            acci[t][v] += A.z * B.x - A.x * B.z;
                            accr[t][v] += A.y * B.y + A.w * B.w;
                            acci[t][v] += A.w * B.y - A.y * B.w;
                        }
                    }
                }
            }
            if (ch + 1 < nchunk) sstore(stg ^ 1);
            __syncthreads();
        }
#pragma unroll
        for (int t = 0; t < NBV; t++) {
            if ((okm >> t) & 1u) {
                f32x16 re, im;
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    re[v] = accr[t][v];
                    im[v] = acci[t][v];
                }
                const int I = SI * SB + bI[t], J = SJ * SB + bJ[t];
                store_block(a, c, fx_bp_index(I, J, nblk), I, J, lane, re, im, nullptr);
            }
        }
    }
}

int xengine_super_blocks(int nblk) { return nblk >= 4 ? 4 : nblk >= 2 ? 2 : 1; }

// Grids of at most this many one-block-tile workgroups stage several K chunks
// per step: at four per CU (32 KB of LDS each) they still run in one round,
// and each workgroup keeps four chunks in flight instead of one (larger grids
// lose more to a second round than they gain).
static constexpr int kSmallGrid = 1024;

template <int PREC>
static hipError_t launch_x_prec(const XArgs& a, hipStream_t s)
{
    const int SB = a.sb > 0 ? a.sb : xengine_super_blocks(a.p.nblk);
    const int npan = (a.p.nblk + SB - 1) / SB;
    const int npairs = npan * (npan + 1) / 2;
    const dim3 grid(a.p.nchan * npairs), block(64 * WAVES);
    constexpr int DEEP = PREC == FX_PREC_F8 ? 2 : 4;    // 32 KB of LDS either way (8 KB for 16x16 blocks)
    switch (SB) {
    case 4:
        // FP16 with one panel pair per channel (at most 64 antennas): every
        // input row is read by exactly one workgroup, and non-temporal loads
        // measured 5.7% faster there (64 x 4096); where panels are shared
        // (larger arrays, smaller tiles) the reuse through the L2 is worth more
        if constexpr (PREC == FX_PREC_F32) {
            fx_xengine_kernel<PREC, 4, 1><<<grid, block, 0, s>>>(a);
        } else if constexpr (PREC == FX_PREC_F16) {
            if (npan == 1) fx_xengine_kernel<PREC, 4, 1, 8, true><<<grid, dim3(64 * 8), 0, s>>>(a);
            else fx_xengine_kernel<PREC, 4, 1, 8><<<grid, dim3(64 * 8), 0, s>>>(a);
        } else {
            fx_xengine_kernel<PREC, 4, 1, 8><<<grid, dim3(64 * 8), 0, s>>>(a);
        }
        break;
    case 2: fx_xengine_kernel<PREC, 2, 1><<<grid, block, 0, s>>>(a); break;
    default:
        if constexpr (PREC == FX_PREC_F32) {
            if (a.p.S <= FX_BLK / 4) {                    // at most 8 signals: two groups per 16x16 MFMA
                // (at least two chunks per step: the partial-sum reduction
                // needs the 2 KB that they stage).  In a small grid each
                // workgroup's K loop is one memory latency per step, and a
                // step of 16 chunks (8 KB) keeps four times the bytes of
                // DEEP chunks in flight (4 antennas x 1024 channels: X-engine
                // 0.86 -> 0.64 ms per 64 subints; 32 chunks measured no faster)
                if ((long long)a.p.nchan * npairs <= kSmallGrid)
                    fx_xengine_kernel<PREC, 1, 16, WAVES, false, false, false, true><<<grid, block, 0, s>>>(a);
                else
                    fx_xengine_kernel<PREC, 1, 2, WAVES, false, false, false, true><<<grid, block, 0, s>>>(a);
                break;
            }
            if (a.p.S <= FX_BLK / 2) {                    // at most 16 signals: 16x16 MFMA blocks
                if ((long long)a.p.nchan * npairs <= kSmallGrid)
                    fx_xengine_kernel<PREC, 1, DEEP, WAVES, false, true><<<grid, block, 0, s>>>(a);
                else
                    fx_xengine_kernel<PREC, 1, 1, WAVES, false, true><<<grid, block, 0, s>>>(a);
                break;
            }
            if (a.p.nblk == 1) {                          // 17 to 32 signals: three 16x16 quarters
                // (two chunks per step in large grids: 16 x 2048 X-engine
                // 0.654 -> 0.645 ms per 16 subints, three A/B pairs)
                if ((long long)a.p.nchan * npairs <= kSmallGrid)
                    fx_xengine_kernel<PREC, 1, DEEP, WAVES, false, false, true><<<grid, block, 0, s>>>(a);
                else
                    fx_xengine_kernel<PREC, 1, 2, WAVES, false, false, true><<<grid, block, 0, s>>>(a);
                break;
            }
        }
        if ((long long)a.p.nchan * npairs <= kSmallGrid) fx_xengine_kernel<PREC, 1, DEEP><<<grid, block, 0, s>>>(a);
        else fx_xengine_kernel<PREC, 1, 1><<<grid, block, 0, s>>>(a);
        break;
    }
    return hipGetLastError();
}

int xengine_chunk_groups(int prec)
{
    return prec == FX_PREC_F8 ? XChunk<FX_PREC_F8>::KC : prec == FX_PREC_F16 ? XChunk<FX_PREC_F16>::KC
                                                                             : XChunk<FX_PREC_F32>::KC;
}


hipError_t launch_xengine(const XArgs& a, int prec, hipStream_t s)
{
    if (a.p.nchan <= 0 || a.ngroups <= 0) return hipSuccess;
    if (a.ngroups % xengine_chunk_groups(prec)) return hipErrorInvalidValue;   // callers pad to whole chunks
    if (a.sb != 0 && a.sb != 1 && a.sb != 2 && a.sb != 4) return hipErrorInvalidValue;
    switch (prec) {
    case FX_PREC_F32: return launch_x_prec<FX_PREC_F32>(a, s);
    case FX_PREC_F16: return launch_x_prec<FX_PREC_F16>(a, s);
    case FX_PREC_F8:  return launch_x_prec<FX_PREC_F8>(a, s);
    case FX_PREC_F32V: {
        const int SB = a.sb > 0 ? a.sb : xengine_super_blocks(a.p.nblk);
        const int npan = (a.p.nblk + SB - 1) / SB;
        const dim3 grid(a.p.nchan * (npan * (npan + 1) / 2)), block(64 * WAVES);
        switch (SB) {
        case 4: fx_xengine_vec_kernel<4><<<grid, block, 0, s>>>(a); break;
        case 2: fx_xengine_vec_kernel<2><<<grid, block, 0, s>>>(a); break;
        default: fx_xengine_vec_kernel<1><<<grid, block, 0, s>>>(a); break;
        }
        return hipGetLastError();
    }
    default: return hipErrorInvalidValue;
    }
}

// counts[j*nant + k] (j <= k) += number of FFT slots where both antennas are
// valid (kernel.c:850; j == k counts autocorrelation FFTs), over the slots one
// X-engine launch reduces.  One thread per (antenna pair j <= k, run of 32
// slots); the per-pair sums of the runs meet in counts[j*nant + k] through
// atomic adds.
__global__ void fx_counts_kernel(int nant, int nslots, int chunk_len, const uint8_t* __restrict__ v,
                                 int* __restrict__ counts)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int pairs = nant * nant;
    const int chunk = (int)(t / pairs), jk = (int)(t - (long long)chunk * pairs);
    const int j = jk / nant, k = jk - j * nant;
    const long long f0 = (long long)chunk * chunk_len;
    if (f0 >= nslots || j > k) return;
    const int f1 = f0 + chunk_len < nslots ? (int)f0 + chunk_len : nslots;
    int n = 0;
#pragma unroll 8
    for (int f = (int)f0; f < f1; f++) n += v[(long long)f * nant + j] & v[(long long)f * nant + k];
    if (n) atomicAdd(counts + jk, n);
}

hipError_t launch_counts(int nant, int nslots, const uint8_t* valid_slots, int* counts, hipStream_t s)
{
    // at most 64 chunks (one atomic each) per counter: the atomics on one
    // address serialise (4 antennas x 40000 slots in chunks of 32 took 48 us)
    const int per = (nslots + 63) / 64;
    const int chunk_len = per <= 32 ? 32 : (per + 31) / 32 * 32;
    const long long n = (long long)nant * nant * ((nslots + chunk_len - 1) / chunk_len);
    if (n <= 0) return hipSuccess;
    fx_counts_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nant, nslots, chunk_len, valid_slots, counts);
    return hipGetLastError();
}

// Block accumulators -> reference layout: cross[b][2l+m][c] for j<k in
// j-major order (kernel.c:831-852) and autos[j][2l+m][c].
// With counts, each visibility is divided by its valid-FFT count where that
// is positive (processNormalize, kernel.c:348-359: (re / n, im / n)).
__global__ void fx_export_kernel(FxPlan p, const float2* __restrict__ vacc, float2* __restrict__ cross,
                                 float2* __restrict__ autos, const int* __restrict__ counts)
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
    if (counts) {
        const int n = counts[j * p.nant + k];
        if (n > 0) {
            const float fn = (float)n;
            v.x = v.x / fn;
            v.y = v.y / fn;
        }
    }
    if (isauto) {
        if (autos) autos[t - ncross] = v;
    } else if (cross) {
        cross[t] = v;
    }
}

hipError_t launch_export(const FxPlan& p, const float2* vacc, float2* cross, float2* autos, const int* counts,
                         hipStream_t s)
{
    const int nbl = p.nant * (p.nant - 1) / 2;
    const long long n = ((long long)nbl + p.nant) * 4 * p.nchan;
    fx_export_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, vacc, cross, autos, counts);
    return hipGetLastError();
}

// FP8 scales from the autocorrelations accumulated so far: per (channel c,
// signal s) the mean parallel-hand auto power P / n over the antenna's valid
// FFTs gives the rms of each spectral component, sqrt(P / 2n), which maps to
// `target` (scale = rms / target, stored value = spectrum / scale).  Signals
// without a valid FFT or power keep their scale.  vacc holds true units (the
// X-engine applies s_i s_j before accumulating), so scales may change between
// X-engine launches.
__global__ void fx_fp8_autoscale_kernel(FxPlan p, const float2* __restrict__ vacc, const int* __restrict__ counts,
                                        float* __restrict__ scale, float* __restrict__ scale_inv, float target)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)p.nchan * p.S) return;
    const int c = (int)(t / p.S), sg = (int)(t - (long long)c * p.S);
    const int n = counts[(sg >> 1) * p.nant + (sg >> 1)];
    const int I = sg / FX_BLK, r = sg % FX_BLK;
    const float P = vacc[((long long)c * p.nbp + fx_bp_index(I, I, p.nblk)) * (FX_BLK * FX_BLK) + r * (FX_BLK + 1)].x;
    if (n <= 0 || !(P > 0.0f) || !(P < 3.0e38f)) return;
    const float rms = sqrtf(P / (2.0f * (float)n));
    const float sc = rms / target;
    if (!(sc > 0.0f) || !(sc < 3.0e38f)) return;
    scale[(long long)c * p.Spad + sg] = sc;
    scale_inv[(long long)sg * p.nchan + c] = 1.0f / sc;
}

hipError_t launch_fp8_autoscale(const FxPlan& p, const float2* vacc, const int* counts, float* scale,
                                float* scale_inv, float target, hipStream_t s)
{
    const long long n = (long long)p.nchan * p.S;
    if (n <= 0) return hipSuccess;
    fx_fp8_autoscale_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, vacc, counts, scale, scale_inv, target);
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

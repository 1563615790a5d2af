// fx_common.h -- shared host/device definitions of the MI355X FX engine.
//
// Data layout in HBM (one context = one subband):
//   packed input   u8   [nant][bytes_per_ant]           2-bit, 4 samples/byte
//                                                        (kernel.c:195-200 layout)
//   delay table    FxDelay [nffts_pad][nant]             FP64 per (FFT, antenna)
//   X input        16-byte rows [nfg/kc][nchan][kc][Spad]  corner-turned spectra,
//                        one row = G consecutive FFTs of one (channel, signal)
//                        (FP32: G = 2, rows (re0, re1, im0, im1)), signal
//                        s = 2*antenna + pol, kc groups per X-engine K chunk;
//                        channels fx_xin_pitch rows apart
//   visibilities   f32x2 [nchan][nbp][32*32]             32x32 blocks of the
//                        Hermitian signal matrix, upper block triangle I <= J
//   counts         int  [nant*nant]                      valid-FFT counts per pair
#pragma once

#include <stdint.h>

#define FX_BLK 32          // X-engine block edge (signals)

// Precision of the F-stage output / X-engine input (same values as difxfx.h).
#ifndef FX_PREC_F32
#define FX_PREC_F32 0
#define FX_PREC_F16 1
#define FX_PREC_F8 2
#define FX_PREC_F32V 3
#endif

struct FxDelay {           // 32 bytes; one per (FFT, antenna)
    double a;              // delay gradient across the FFT (s)
    double b;              // delay at FFT start (s)
    double fd;             // fractional delay (s) after the edge clamps
    int offset;            // first sample of the FFT window in the stream
    int valid;             // kernel.c:752-756
};

struct FxPlan {
    int nant, nchan, F, lg;
    int stride, sub, nstep, nfstep;  // rotator factorisation
    int S, Spad, nblk, nbp;          // signals, padded signals, blocks, block pairs
    int G;                           // FFTs per corner-turn group
    int fracLO;
    int iscomplex;                   // 0: real 2-bit input (F = 2*nchan); 1: complex 8-bit (F = nchan)
    double Ts, lo, bw, lofrac;       // lofrac = lo - (int)lo
};

// Rotator tables of one (FFT, antenna): fringe sub [sub], fringe step
// [nstep], fractional sub [stride], fractional step [nfstep], then the fringe
// steps again in the order the F-stage's scalar loads read them
// (step b + q*r at b*(nstep/q) + r, q set by the radix plan).
static inline __host__ __device__ int fx_rtab_len(const FxPlan& p)
{
    return p.sub + 2 * p.nstep + p.stride + p.nfstep;
}

// Rows between consecutive channels of one K chunk of the X input: kc*Spad
// plus five 128-byte lines.  The F-stage writes one 16-byte row per (channel,
// signal), and eight signals' rows (eight workgroups) fill a line.  At a
// power-of-two pitch the rows one workgroup writes fall into few L2 sets, and
// lines left the L2 before their eight signals met there: the F-stage wrote
// 1.14x to 3.41x the spectra's bytes over the four shapes measured (64 and
// 256 antennas, FP16 and FP8 rows), 1.04x to 2.18x with one line of padding
// and 1.03x to 1.50x with five.
static inline __host__ __device__ long long fx_xin_pitch(int kc, int Spad)
{
    return (long long)kc * Spad + 40;
}

// Upper-triangle block-pair index of (I, J), I <= J < nblk, row-major.
static inline __host__ __device__ int fx_bp_index(int I, int J, int nblk)
{
    return I * nblk - I * (I - 1) / 2 + (J - I);
}

static inline int fx_ilog2_exact(int n)
{
    int lg = 0;
    if (n <= 0) return -1;
    while ((1 << lg) < n) lg++;
    return ((1 << lg) == n) ? lg : -1;
}

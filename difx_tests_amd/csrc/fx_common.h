// fx_common.h -- shared host/device definitions of the MI355X FX engine.
//
// Data layout in HBM (one context = one subband):
//   packed input   u8   [nant][bytes_per_ant]           2-bit, 4 samples/byte
//                                                        (kernel.c:195-200 layout)
//   delay table    FxDelay [nffts_pad][nant]             FP64 per (FFT, antenna)
//   X input        T    [nchan][nfg][Spad][G]            corner-turned spectra,
//                        (FP32: G = 2, rows (re0, re1, im0, im1))
//                        signal s = 2*antenna + pol, G consecutive FFTs per group
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

// fx_kernels.h -- launch interface between the context code and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include "difxfx.h"
#include "fx_common.h"

namespace fx {

// Plan of fx_params (checks them; the error is left in fx_last_error) and the
// error message of a failing call (returns code).
int plan_from_params(const fx_params* prm, FxPlan* p);
int set_error(int code, const char* msg);

struct FStageArgs {
    FxPlan p;
    const uint8_t* data;          // packed input, antenna a at data + a*data_stride
    long long data_stride;
    const FxDelay* delays;        // [nffts][nant] for this call
    const float2* rtab;           // [nffts][nant][sub + nstep + stride + nfstep] rotator tables
    const float2* tw;             // [F] exp(-2 pi i m / F)
    const float* subchf;          // [stride] f32 channel frequencies (kernel.c:164)
    const float* stepchf;         // [nfstep] (kernel.c:165)
    void* out;                    // X input, [nfg/kc][nchan][kc][Spad] rows of G FFTs, channel pitch fx_xin_pitch
    int kc;                       // groups per X-engine K chunk
    int fg0;                      // first group written by this launch
    int ngroups;                  // groups in this launch
    const float* scale_inv;       // FP8: [Spad][nchan] inverse scales (channel fastest: coalesced)
    float2* tap;                  // optional FP32 spectra [nffts][nant][2][nchan]
    float* ltap;                  // optional decoded 2-bit levels [nffts][nant][2][F] of the
                                  // valid FFTs (real input; only with tap)
    void* stage;                  // F = 32768: row staging, fx_fstage_stage_bytes(lg, prec, nant, ngroups) bytes
    int plan;                     // radix plan, 0 .. fstage_plans(lg, iscomplex) - 1
    int tile_ant;                 // antennas per work tile (set by launch_fstage)
};

struct XArgs {
    FxPlan p;
    const void* xin;              // [nfg/KC][nchan][KC][Spad] rows of G FFTs, channel pitch fx_xin_pitch
    int fg0, ngroups;             // group range reduced by this launch (multiples of KC)
    float2* vacc;                 // [nchan][nbp][32*32]
    int accumulate;               // 0: overwrite, 1: add
    const float* scale;           // FP8: [nchan][Spad] scales
    int sb;                       // super-block size 1, 2 or 4; 0 picks by the block count
};

// F-stage: LG = log2(F); prec = FX_PREC_*; G is fixed per precision.
int fx_fstage_group_size(int prec);
// origin_dev: device int, global FFT index of the data buffer's first FFT
// (read by the kernel, so a recorded graph serves every subint)
hipError_t launch_delay(const FxPlan& p, const double* delays, const double* fileoff, const int* iteration,
                        int numffts, int nffts_pad, int maxoffset, const int* origin_dev, FxDelay* out,
                        hipStream_t s);
// Delay entries, FFT-slot validity bytes and rotator tables of one call in one
// launch.  With origin_dev null: iteration is a device FFT-index list, or
// null for it0 + k, and the buffer origin is `origin`.  Otherwise origin_dev
// points to device {origin, first index, list flag}: FFT k is iteration[k]
// when the flag is set, else first index + k (a recorded graph reads every
// subint's values from device memory).  valid_slots: [nffts_pad][nant] bytes of the call's slots.
// step_q: order of the tables' second copy of the fringe steps
// (fstage_step_q; 0 leaves it unwritten).
hipError_t launch_delay_rotators(const FxPlan& p, const double* delays, const double* fileoff, const int* iteration,
                                 int it0, int numffts, int nffts_pad, int maxoffset, const int* origin_dev, int origin,
                                 const float* subchf, const float* stepchf, int step_q, FxDelay* dtab, float2* rtab,
                                 uint8_t* valid_slots, hipStream_t s);
// The delay / rotator kernel's function as graph kernel nodes record it, and
// the positions of its first-FFT-index and origin arguments among its
// kDelayNArgs arguments (a recorded graph's launches update those two).
const void* delay_rot_kernel_fn();
constexpr int kDelayArgIt0 = 4, kDelayArgOrigin = 9, kDelayNArgs = 16;
// The F-stage kernel of radix plan `plan` reads its pass-0 fringe steps b + q*r
// (r < R0) with scalar loads from the tables' second copy of the steps; returns
// that q, or 0 when the plan reads them from the LDS.
int fstage_step_q(int lg, int plan, int iscomplex);
hipError_t launch_fstage(const FStageArgs& a, int prec, hipStream_t s);
// Device bytes of FStageArgs::stage for a launch of ngroups groups (0 where
// the F-stage stages nothing, F < 32768)
size_t fx_fstage_stage_bytes(int lg, int prec, int nant, int ngroups);
bool fstage_supported(int lg);
bool fstage_supported_complex(int lg);   // complex 8-bit input, F = nchan
int fstage_plans(int lg, int iscomplex);
// Batched forward C2C FFT of length 2^lg (natural order in/out, unnormalised).
hipError_t launch_fft_c2c(const float2* in, float2* out, const float2* tw, int lg, int batch, hipStream_t s);

// X-engine; ngroups must be a multiple of xengine_chunk_groups(prec).
int xengine_chunk_groups(int prec);
int xengine_super_blocks(int nblk);
hipError_t launch_xengine(const XArgs& a, int prec, hipStream_t s);
// counts[j*nant + k] (j <= k) += FFT slots of valid_slots[nslots][nant] where
// both antennas are valid
hipError_t launch_counts(int nant, int nslots, const uint8_t* valid_slots, int* counts, hipStream_t s);
// counts (may be null): [nant*nant] valid-FFT counts; when given, the export
// is normalised by them.
hipError_t launch_export(const FxPlan& p, const float2* vacc, float2* cross, float2* autos, const int* counts,
                         hipStream_t s);
hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s);
// FP8 scales [nchan][Spad] and inverse scales [Spad][nchan] from the
// accumulated autocorrelations and valid-FFT counts (rms per component ->
// target).
hipError_t launch_fp8_autoscale(const FxPlan& p, const float2* vacc, const int* counts, float* scale,
                                float* scale_inv, float target, hipStream_t s);

}  // namespace fx

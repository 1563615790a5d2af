// fx_single.h -- shared state between the kernel.h orchestration entry points
// and the single-shot (one antenna, one FFT) GPU entry points.
#pragma once

#include <stdint.h>

#include "fx_common.h"
#include "kernel.h"

// FFT length set by the last init_FxKernel (0 before it).  dofft has no
// length parameter in kernel.h:76; the reference takes it from the FFT spec
// that allocKernel built for the same nchan.
int fxk_fftchannels(void);
// Bandwidth of the same init_FxKernel: fracSampleCorrect (kernel.h:78) takes
// the channel frequencies from the FFT spec built for that bandwidth.
double fxk_bandwidth(void);
void fxk_set_geometry(int fftchannels, double bw);
// Each antenna's unpacked / channelised / conjugated buffers of its last
// valid FFT (last[j], has[j] != 0), formed on the device from the packed
// input d_data (antenna a at d_data + a*stride) and copied to the caller's
// host arrays (kernel.c:725-800).
void fxk_fill_antenna_buffers(const uint8_t* d_data, long long stride, const FxDelay* last, const int* has,
                              int nant, int F, int nchan, int sub, int stridesize, double lo, int fracLO, double bw,
                              cf32*** unpacked, cf32*** channelised, cf32*** conjchannels);

// fx_single.h -- shared state between the kernel.h orchestration entry points
// and the single-shot (one antenna, one FFT) GPU entry points.
#pragma once

// FFT length set by the last init_FxKernel (0 before it).  dofft has no
// length parameter in kernel.h:76; the reference takes it from the FFT spec
// that allocKernel built for the same nchan.
int fxk_fftchannels(void);
// Bandwidth of the same init_FxKernel: fracSampleCorrect (kernel.h:78) takes
// the channel frequencies from the FFT spec built for that bandwidth.
double fxk_bandwidth(void);
void fxk_set_geometry(int fftchannels, double bw);

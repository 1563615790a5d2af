/*
 * vectordefs.h -- the type names of the reference's vectordefs.h
 * (vectordefs.h:11-16, 201) without Intel IPP.
 *
 * The reference maps u8/f32/f64/cf32/cf64 onto IPP types with macros; here
 * they are plain C types with the same size and layout (Ipp32fc is
 * {float re; float im;}, which callers rely on: main.c:444 reads .im).
 * The IPP vector primitives themselves are not part of this interface: the
 * FX engine behind kernel.h runs on the GPU.
 */
#ifndef VECTORDEFS_H
#define VECTORDEFS_H

#include <math.h>

#define TWO_PI 2 * 3.14159265358979323846  /* vectordefs.h:8, IPP_PI */
#define TINY 0.00000000001                 /* vectordefs.h:9 */

typedef unsigned char u8;
typedef float f32;
typedef double f64;
typedef struct { float re; float im; } cf32;
typedef struct { double re; double im; } cf64;

/* FFT specification handle (vectordefs.h:65, IppsFFTSpec_C_32fc): opaque. */
typedef struct vecFFTSpecC_cf32_s vecFFTSpecC_cf32;

#endif

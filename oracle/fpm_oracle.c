/*
 * fpm_oracle.c -- CPU oracle of the fused FPM reconstruction.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of libfpm_amd; only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
 * it.  It restates the algorithm of THUHoloLab/Fused-Fourier-Ptycho's
 * cuFPM_v3 (cuFPM_v3/cuFPM_pure_v3.cu:44-228, cuFPM_v3/cuda/kernels.cuh) on
 * the CPU, in FP64 (the per-step gradient oracle) and in FP32 with v3's
 * operation order and a deterministic stitch (the trajectory oracle and the
 * CPU baseline).  See oracle_impl.h for the per-function citations.
 *
 * Pins (all in tests/): the FFT against numpy.fft (pocketfft, FP64) and a
 * brute-force DFT; the crop/shift indices against numpy.fft.fftshift and a
 * closed form; the whole gradient against central finite differences of the
 * loss (G_O = r^2 grad_O L, G_P = n^2 r^4 grad_P L for N = n r); the update
 * against an elementwise scalar restatement and the pupil-amplitude invariant.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX f64
#define M_SQRT sqrt
#define M_COPYSIGN copysign
#define M_FABS fabs
#define M_ATAN2 atan2
#define M_COS cos
#define M_SIN sin
#define M_FMAX fmax
#define M_FMIN fmin
#include "oracle_impl.h"
#undef REAL
#undef SFX
#undef M_SQRT
#undef M_COPYSIGN
#undef M_FABS
#undef M_ATAN2
#undef M_COS
#undef M_SIN
#undef M_FMAX
#undef M_FMIN

#define REAL float
#define SFX f32
#define M_SQRT sqrtf
#define M_COPYSIGN copysignf
#define M_FABS fabsf
#define M_ATAN2 atan2f
#define M_COS cosf
#define M_SIN sinf
#define M_FMAX fmaxf
#define M_FMIN fminf
#include "oracle_impl.h"

/* uint16 raw stack -> float amplitudes, as codes_v2/main_run_FPM.m:43-45 does
 * in single precision: (I - min) / max(I - min), then sqrt. */
void oracle_prepare_amplitude(const uint16_t* raw, long count, float* out) {
    uint16_t lo = 0xffff, hi = 0;
    for (long k = 0; k < count; ++k) {
        if (raw[k] < lo) lo = raw[k];
        if (raw[k] > hi) hi = raw[k];
    }
    const float range = (float)(hi - lo);
    for (long k = 0; k < count; ++k) out[k] = sqrtf((float)(raw[k] - lo) / range);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

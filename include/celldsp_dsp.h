/*
 * celldsp_dsp.h -- DSP task commands and the spe* host-pointer wrappers.
 *
 * Drop-in for cellspe-tasklib-dsp.h:27-95.  Each spe* call builds a
 * CellDSPTask_t and runs it through celldsp_executeTask (or the continuable
 * calls) exactly like cellspe-tasklib-dsp.cpp:233-359; the task executes on a
 * gfx950 unit.  Pointers are host pointers owned by the caller; lengths are in
 * elements as in the reference.  Return value: vecNoErr (0) on success,
 * !vecNoErr (1) otherwise.
 */
#ifndef CELLDSP_DSP_H
#define CELLDSP_DSP_H

#include "celldsp_tasklib.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cellspe-tasklib-dsp.h:27-46 */
enum CellDspCommandEnum {
    CELLDSP_TASK_NOP = 0,
    CELLDSP_TASK_MAC,            /* out[] += in1[] * in2[]  (complex, no conjugate) */
    CELLDSP_TASK_MULC,           /* out[] = in1[] * in2[]   (complex)               */
    CELLDSP_TASK_ADDC,           /* out[] = in1[] + in2[]   (complex)               */
    CELLDSP_TASK_IOTEST,         /* out[0..3] = {0.1, 2.3, 4.5, 6.7}                 */
    CELLDSP_TASK_IOBENCH,        /* outputs round-trip through device memory         */
    CELLDSP_TASK_FFT_INIT,       /* reserve a unit for a continuable FFT of nfft      */
    CELLDSP_TASK_FFT_FREE,
    CELLDSP_TASK_FFT_EXEC,       /* forward complex FFT of the initialised size       */
    CELLDSP_TASK_FFT_EXEC_DYN,   /* FFT of subcontext.FFT.nfft, re-planned on change  */
    CELLDSP_TASK_SINCOS,         /* (cos, sin) of a phase vector                      */
    CELLDSP_TASK_SINCOS_NORM,    /* same for phases in turns (1.0 = 2 pi)             */
    CELLDSP_TASK_SINCOS_GENERTR, /* oscillator (cos, sin)(start + i*delta)            */
    CELLDSP_TASK_UNPACK,         /* 2-bit words -> float [nchan][nsamp]               */
    CELLDSP_TASK_RUNALL,
    CELLDSP_TASK_DMASIMULATION,
    CELLDSP_TASK_DIFXPROCESS,    /* FFT, sideband, fractional-sample rotation, auto  */
    CELLDSP_TASK_SIGTERM
};

typedef struct { float re; float im; } fc32;
typedef fc32 cf32;

int speNOP(void);
/* cellspe-tasklib-dsp.cpp:111-124: IOTEST into a static buffer, prints the
 * four values written back; vecNoErr when the task ran */
int speIOCheck(void);
/* not in the reference: IOTEST into the caller's 4 floats (no printing) */
int speIOCheck_out(float* out4);
int speAddProduct_32fc(fc32* src, fc32* src2, fc32* accumulator, int length);
int speMul_32fc(fc32* src, fc32* src2, fc32* dest, int length);
int speAdd_32fc(fc32* src, fc32* src2, fc32* dest, int length);
int speFFT_init(int fft_len, int* speID);
int speFFT_exec(void* in, void* out, int fft_len, int* speID);
int speFFT_destroy(int* speID);
int speFFT_singleexec(void* in, void* out, int fft_len);
int speDIFX_process(void* sincosrotdone, void* complexfracmult, void* fftoutJ, void* conjfftoutJ,
                    void* autocorr, int nfft, int twicenumchannels, int useUpperOrLower);
int speSinCosGenerator(float startphase, float increment, fc32* cos_sin_out, int numiter);
int speSinCosAccurate(float* argument, cf32* cos_sin_out, int argumentcount);
int speSinCosAccurateNorm(float* normalizedargument, cf32* cos_sin_out, int argumentcount);
int speUnpack2bit(const unsigned int* words, float* out, int nsamp, int nchan);

#ifdef __cplusplus
}
#endif
#endif

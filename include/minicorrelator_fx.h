/*
 * minicorrelator_fx.h -- the minicorrelator's control-block drop-in.
 *
 * Replaces the PPU program minicorrelator_ppu.c:90-262, which fills one
 * control_block per station (minicorrelator.h:107-119) and starts one SPE per
 * station.  minicorr_run() takes the same control blocks and computes the same
 * products on one MI355X: 2-bit VDIF-convention unpack, N_FFT-point real FFT,
 * optional per-channel phase rotation, conjugate X-engine for every baseline
 * and auto-correlation, integration dumps every ffts_to_integrate spectra.
 *
 * Output layout is the reference's: for every dump block b the product buffer
 * holds PPU_CORRELBUF_BYTES = [Re 0..N_FFT-1][Im 0..N_FFT-1] floats at byte
 * offset b * PPU_CORRELBUF_BYTES (store_data, minicorrelator_spu.c:97-102).
 * Bins 0..N_FFT/2-1 are the computed channels, bin N_FFT/2 (Nyquist) is 0,
 * and bins N_FFT/2+1..N_FFT-1 carry the Hermitian mirror conj(V(N_FFT-k)) that
 * a complex N_FFT-point transform of real data would produce.
 *
 * Baselines are numbered canonically: (0,1)=0, (0,2)=1, ..., (0,5)=4,
 * (1,2)=5, ..., (4,5)=14 (the reference's spreadout table,
 * minicorrelator_spu.c:52-53,842-857, leaves baselines out).
 */
#ifndef MINICORRELATOR_FX_H
#define MINICORRELATOR_FX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* minicorrelator.h:84-104 */
#define NUM_SPE_THREADS 6
#define NUM_BASELINES (((NUM_SPE_THREADS - 1) * NUM_SPE_THREADS) / 2)
#define N_FFT 1024
#define SAMPLES_PER_BYTE 4
#define PPU_CORRELBUF_BYTES (N_FFT * 2 * sizeof(float))

#ifndef FX_ADDR64_DEFINED
#define FX_ADDR64_DEFINED
/* 32/64-bit address union (minicorrelator.h:66-69, cell_spectrometer.h:96-100) */
typedef union {
    unsigned long long ull;
    unsigned int ui[2];
    void* p;
} addr64;
#endif

/* minicorrelator.h:107-119 */
typedef struct _control_block {
    addr64 rawdata_src __attribute__((aligned(128)));   /* 2-bit samples, ffts_total*N_FFT/4 bytes */
    addr64 autocorrelation_out;                           /* this station's auto dumps */
    addr64 baselines_out[15];                             /* every baseline's dumps (same in all blocks) */
    addr64 fmultipliers_src;                              /* optional: N_FFT/2 floats, phase in turns */
    addr64 syncline;                                      /* unused (SPE barrier) */
    int spe_num;                                          /* station index */
    int num_spes;                                         /* number of stations */
    unsigned int ffts_total;                              /* spectra to process */
    unsigned int ffts_to_integrate;                       /* spectra per dump */
    addr64 spe_ls_listOnPPU;                              /* unused (SPE local stores) */
    addr64 spe_sig_listOnPPU;                             /* unused (SPE signals) */
} control_block __attribute__((aligned(128)));

typedef struct {
    double device_ms;          /* device time of the whole run (hipEvents) */
    unsigned long long spectra;
    unsigned int dumps;
} minicorr_stats;

/* Run the correlator over nstations control blocks (cb[s].spe_num == s).
 * 2 <= nstations <= NUM_SPE_THREADS.  Returns 0 on success, -1 on error
 * (message via fx_last_error()).  stats may be NULL. */
int minicorr_run(const control_block* cb, int nstations, minicorr_stats* stats);

#ifdef __cplusplus
}
#endif
#endif

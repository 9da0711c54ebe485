/*
 * fx_oracle.h -- CPU restatement of the cellspe-tasklib F-X chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in cellspe-tasklib_amd/: only tests/, the smoke() entry of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  No
 * product code links it, calls it or falls back to it.
 *
 * Every routine follows the reference's GENERIC (non-SPU) semantics and cites
 * the reference file:line it restates.  FP32 routines reproduce the
 * reference's arithmetic order; FP64 routines are the high-precision oracle
 * the reduced-precision X-engines are measured against.
 */
#ifndef FX_ORACLE_H
#define FX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float re, im; } fxo_cf32;   /* types_generic.h:44-45 (cf32) */
typedef struct { double re, im; } fxo_cf64;

/* ------------------------------------------------------------------ */
/* Synthetic signals: counter-based RNG, quantizer, packer              */
/* ------------------------------------------------------------------ */
uint64_t fxo_mix64(uint64_t x);
double   fxo_uniform(uint64_t key, uint64_t ctr);         /* in (0,1) */
double   fxo_gauss(uint64_t key, uint64_t ctr);            /* N(0,1), Box-Muller FP64 */
uint64_t fxo_stream_key(uint64_t seed, uint64_t stream);
int      fxo_quantize2(double x, double thresh);           /* offset-binary code 0..3 */
void     fxo_pack2b(const uint8_t* codes, size_t n, uint32_t* words);

/* One station stream: x[n] = rho*common[n-delay] + sqrt(1-rho^2)*own[n]
 *                          + tone_amp*cos(2*pi*tone_freq*n)
 * quantized at {-thresh,0,+thresh}, packed 16 codes per LE u32 word.
 * common[] uses stream key (seed, 0xC0); own[] uses (seed, 0x100+station). */
void fxo_gen_station(uint64_t seed, int station, double rho, long delay,
                     double tone_amp, double tone_freq_norm, double thresh,
                     size_t nsamp, uint32_t* words);

/* Interleaved multi-channel stream (VDIF nchan x 2-bit layout): channel c of
 * time sample t at bit 2*(t%spw*nchan + c) of word t/spw, spw = 16/nchan.
 * Channel c is own-noise N(0,1) plus tone_amp[c]*cos(2*pi*tone_freq[c]*t). */
void fxo_gen_multichannel(uint64_t seed, int nchan, const double* tone_amp,
                          const double* tone_freq_norm, double thresh,
                          size_t ntime, uint32_t* words);

/* Channelized complex voltages, Re and Im independently 2-bit quantized from
 * rho*common + sqrt(1-rho^2)*own Gaussians; returns codes [c][t][i] packed
 * as re | im<<2 per byte. */
void fxo_gen_channelized_codes(uint64_t seed, int nchan, long ntime, int ninp,
                               double rho, double thresh, uint8_t* codes);

/* ------------------------------------------------------------------ */
/* a1: 2-bit unpack                                                      */
/* ------------------------------------------------------------------ */
extern const float fxo_lut2bit[4];   /* format_vdif.c:70 {-HiMag,-1,1,HiMag} */
void fxo_unpack2b_1ch(const uint32_t* words, size_t nsamp, float* out);
void fxo_unpack2b_nch(const uint32_t* words, size_t ntime, int nchan, float* out);

/* ------------------------------------------------------------------ */
/* a2/a3: radix-2 FFT                                                    */
/* ------------------------------------------------------------------ */
void fxo_twiddle_r2(int n, fxo_cf32* W);                       /* n/4 entries */
void fxo_fft_r2_c2c(fxo_cf32* data, int n, const fxo_cf32* W); /* in place, forward */
void fxo_rfft_r2(const float* x, int n, const fxo_cf32* W, fxo_cf32* X, fxo_cf32* work);
void fxo_fft_f64(fxo_cf64* data, int n);                       /* forward, exact twiddles */
void fxo_dft_f64(const fxo_cf64* in, int n, fxo_cf64* out);    /* brute force O(n^2) */

/* ------------------------------------------------------------------ */
/* a4: FP64 phase model + FP32 rotation                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    double delay0;       /* samples */
    double delay_rate;   /* samples per spectrum */
    double fringe0;      /* radians */
    double fringe_rate;  /* radians per spectrum */
} fxo_phase_model;

double fxo_phase(const fxo_phase_model* m, int k, int n, long t);
void   fxo_rotate(fxo_cf32* X, int nchan, int n, const fxo_phase_model* m, long t);
void   fxo_rotate_f64(fxo_cf64* X, int nchan, int n, const fxo_phase_model* m, long t);

/* ------------------------------------------------------------------ */
/* a6: complex MAC primitives and X-engine                               */
/* ------------------------------------------------------------------ */
void   fxo_mul(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* out, int len);
void   fxo_addproduct(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* acc, int len);
void   fxo_addproduct_conj(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* acc, int len);
size_t fxo_prod_index(int i, int j, int ninp);
size_t fxo_nprod(int ninp);

/* X-engine over channelized FP32 data X[c][t][i]; vis[p][c] += sum_t X_i conj(X_j). */
void fxo_xengine_f32(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf32* vis);
void fxo_xengine_f32_mt(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf32* vis, int nthreads);
void fxo_xengine_f64(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf64* vis);
/* X[t][i][c] channelized data, reference loop order, time-sliced workers */
void fxo_xcorr_tic_f32(const fxo_cf32* X, int ninp, int nchan, long ntime, fxo_cf32* vis, int nthreads);

/* Full F-X chain, FP32 (generic-C semantics): words[i] = packed stream of
 * input i (ntime*n samples), vis[p][c] with c = n/2 channels.  pm may be
 * NULL (no rotation).  nthreads > 1 splits the time axis (time-sliced
 * TaskDispatcher model) and sums per-thread accumulators in thread order. */
void fxo_corr_f32(const uint32_t* const* words, int ninp, int n, long ntime,
                  const fxo_phase_model* pm, fxo_cf32* vis, int nthreads);
/* Same chain in FP64 (FP64 FFT, FP64 rotation, FP64 MAC): the FP64 oracle. */
void fxo_corr_f64(const uint32_t* const* words, int ninp, int n, long ntime,
                  const fxo_phase_model* pm, fxo_cf64* vis, int nthreads);
/* F-stage only: channelized FP32 output X[c][t][i] for t in [t0, t0+nt). */
void fxo_fstage_f32(const uint32_t* const* words, int ninp, int n, long t0, long nt,
                    const fxo_phase_model* pm, fxo_cf32* X);

/* ------------------------------------------------------------------ */
/* a8: spectrometer                                                      */
/* ------------------------------------------------------------------ */
enum {
    FXO_WIN_NONE = 0, FXO_WIN_SIN2 = 1, FXO_WIN_COSINE = 2, FXO_WIN_COSINE2 = 3,
    FXO_WIN_HAMMING = 4, FXO_WIN_BLACKMAN = 5, FXO_WIN_HANN = 6
};
void fxo_window(int kind, int n, float* w);
/* generalized Blackman (alpha+1)/2 - cos(x)/2 - (alpha/2) cos(2x),
 * x = 2 pi i/(n-1); the IPP core calls it with alpha = 0.5
 * (TaskCoreIPP.cpp:591-594), alpha = -0.16 is the classical Blackman */
void fxo_window_blackman(double alpha, int n, float* w);
/* nchan interleaved 2-bit channels (VDIF layout), n-point real FFT per
 * channel, |X|^2 accumulated over integrate_ffts FFTs then scaled by
 * 1/integrate_ffts.  out[s][c][n/2] for s in [0, ntime/(n*integrate_ffts)). */
void fxo_spectrometer_f32(const uint32_t* words, size_t ntime, int nchan, int n,
                          int integrate_ffts, int window_kind, float* out, int nthreads);
/* keep_nyquist: rows of n/2+1 bins (fft_ssb_points) instead of n/2 */
size_t fxo_spectrometer_ex_f32(const uint32_t* words, size_t ntime, int nchan, int n, int overlap,
                               int integrate_ffts, int window_kind, int continuous, int keep_nyquist,
                               float* out, int nthreads);
void fxo_spectrometer_f64(const uint32_t* words, size_t ntime, int nchan, int n,
                          int integrate_ffts, int window_kind, double* out);

/* ------------------------------------------------------------------ */
/* a5: phase-cal extraction (FP64)                                      */
/* ------------------------------------------------------------------ */
enum { FXO_PCAL_TRIVIAL = 0, FXO_PCAL_IMPLICIT = 1, FXO_PCAL_SHIFTING = 2 };
typedef struct {
    int    kind;
    double fs_hz, spacing_hz, offset_hz;
    int    nbins;      /* folding length */
    int    ntones;
    int    rotatorlen; /* Shifting only */
    int    step, offset_bin;
} fxo_pcal_cfg;
long long fxo_gcd(long a, long b);
void fxo_pcal_config(double bandwidth_hz, double spacing_hz, int offset_hz, fxo_pcal_cfg* cfg);
/* fold state: acc has 2*nbins complex (real kinds use .re only); index = fold position */
void fxo_pcal_extract(const fxo_pcal_cfg* cfg, const float* x, size_t len,
                      size_t sampleoffset, fxo_cf64* acc);
void fxo_pcal_final(const fxo_pcal_cfg* cfg, const fxo_cf64* acc, fxo_cf64* tones);
void fxo_pcal_reference(const fxo_pcal_cfg* cfg, const float* x, size_t len,
                        size_t sampleoffset, fxo_cf64* tones);

/* The IPP spectrometer core's own phase-cal (TaskCoreIPP.cpp:128-136,
 * 413-418, 633-693; parameters swspectrometer.cpp:178-196): each FFT input
 * of `points` samples is multiplied by the rotator e^{i dphi r}, r = n mod
 * rotatorlen, dphi = -2 pi offset/fs, and folded into tonebins bins:
 * acc[r mod tonebins] += x[n] e^{i dphi r} for r < pulses*tonebins,
 * pulses = rotatorlen / tonebins.  FP64 throughout. */
void fxo_tcpcal_fold(const float* x, int points, int tonebins, int rotatorlen, double dphi,
                     fxo_cf64* acc);

#ifdef __cplusplus
}
#endif
#endif

/*
 * cellspe_fx.h -- C-ABI of the MI355X (gfx950) F-X engine.
 *
 * Device-pointer API of the hot path: 2-bit unpack, FFT, FP64 phase model
 * rotation, phase-cal extraction, the X-engine (FP32 / FP16 / FP8 MFMA) and
 * the spectrometer power accumulation.  The reference's own drop-in surfaces
 * (celldsp_* task runtime, spe* DSP wrappers, minicorrelator control block,
 * spectrometer control block) are declared in celldsp_tasklib.h,
 * celldsp_dsp.h, minicorrelator_fx.h and spectrometer_fx.h and are built on
 * these calls.
 *
 * All calls are plain C: pointers, sizes and integer codes.  Work is enqueued
 * on the context's HIP stream; calls return once the work is enqueued unless
 * stated otherwise.  Every call returns FX_OK (0) or a negative FX_ERR_* code
 * and leaves a message for fx_last_error().
 */
#ifndef CELLSPE_FX_H
#define CELLSPE_FX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
enum {
    FX_OK = 0,
    FX_ERR_ARG = -1,         /* invalid argument / unsupported size       */
    FX_ERR_HIP = -2,         /* HIP runtime error                          */
    FX_ERR_NOMEM = -3,       /* device allocation failed                   */
    FX_ERR_UNSUPPORTED = -4, /* valid request this build does not support  */
    FX_ERR_STATE = -5        /* call out of order (finalized, no device)   */
};
const char* fx_last_error(void);

/* ---- types --------------------------------------------------------------- */
typedef struct { float re, im; } fx_cf32;   /* types_generic.h:44-45 cf32 */
typedef struct { double re, im; } fx_cf64;

/* FP64 phase model per input: phi(k,t) = 2 pi k tau(t)/N + fringe(t),
 * tau(t) = delay0 + delay_rate t [samples], fringe(t) = fringe0 + fringe_rate t
 * [rad], t = spectrum index.  Reduced mod 2 pi in FP64, applied in FP32. */
typedef struct {
    double delay0, delay_rate, fringe0, fringe_rate;
} fx_phase_model;

/* X-engine input precisions */
enum {
    FX_PREC_F32 = 0,     /* plain FP32 MFMA (v_mfma_f32_32x32x2_f32)              */
    FX_PREC_F16 = 1,     /* FP16 in, FP32 accumulate (v_mfma_f32_32x32x16_f16)     */
    FX_PREC_FP8 = 2,      /* FP8-e4m3 in, FP32 accumulate, per-channel/input scales */
    FX_PREC_FP8_2BIT = 3, /* exact FP8 decomposition of 2-bit {+-1,+-H} voltages    */
    FX_PREC_FP4_2BIT = 4  /* the same exact decomposition on FP4-e2m1 MFMA          */
};

/* recorded-data frame formats of the front end */
enum {
    FX_FMT_VDIF = 0,   /* VDIF: 32-byte header (16 legacy), payload of 8-byte multiples */
    FX_FMT_MARK5B = 1, /* Mark5B: 16-byte header, 10000-byte payload, 10016-byte frames */
    FX_FMT_RAW = 2     /* headerless packed 2-bit samples (this library's layout) */
};

/* spectrometer windows: SIN2 sin^2(pi i/(n-1)) (spectrometer_spu.c:297-315);
 * COSINE, COSINE2, HAMMING, BLACKMAN (alpha 0.5) and HANN as the IPP core's
 * generate_windowfunction (TaskCoreIPP.cpp:569-605) */
enum {
    FX_WIN_NONE = 0, FX_WIN_SIN2 = 1, FX_WIN_COSINE = 2, FX_WIN_COSINE2 = 3,
    FX_WIN_HAMMING = 4, FX_WIN_BLACKMAN = 5, FX_WIN_HANN = 6
};

/* ---- context ------------------------------------------------------------- */
typedef struct fx_ctx fx_ctx;
int   fx_ctx_create(int device, fx_ctx** ctx);
int   fx_ctx_destroy(fx_ctx* ctx);
int   fx_ctx_sync(fx_ctx* ctx);
void* fx_ctx_stream(fx_ctx* ctx);           /* hipStream_t of the context */
int   fx_device_count(void);
int   fx_device_name(int device, char* buf, int len);

/* ---- memory and timing helpers ------------------------------------------ */
int fx_malloc(fx_ctx* ctx, size_t bytes, void** dptr);
int fx_free(fx_ctx* ctx, void* dptr);
int fx_memcpy_h2d(fx_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int fx_memcpy_d2h(fx_ctx* ctx, void* dst, const void* src, size_t bytes);   /* synchronous */
int fx_memcpy_d2d(fx_ctx* ctx, void* dst, const void* src, size_t bytes);   /* async */
int fx_memset(fx_ctx* ctx, void* dst, int value, size_t bytes);             /* async; a fill kernel when dst
                                                                               and bytes are 4-byte aligned */
int fx_event_create(fx_ctx* ctx, void** ev);
int fx_event_record(fx_ctx* ctx, void* ev);
int fx_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms);          /* syncs stop */
int fx_event_destroy(void* ev);

/* ---- synthetic signal generators (benchmarks and tests) ----------------- */
/* One 2-bit station stream of nsamp samples (see fx_oracle for the model). */
int fx_synth_station(fx_ctx* ctx, uint64_t seed, int station, double rho, long delay,
                     double tone_amp, double tone_freq_norm, double thresh, size_t nsamp,
                     uint32_t* d_words);
/* nchan-interleaved 2-bit stream (VDIF layout), per-channel tone (host arrays). */
int fx_synth_multichannel(fx_ctx* ctx, uint64_t seed, int nchan, const double* tone_amp,
                          const double* tone_freq_norm, double thresh, size_t ntime,
                          uint32_t* d_words);
/* 2-bit channelized voltages, code = re | im << 2, stored codes[c][t][i]
 * (input_major = 0) or codes[c][i][t] (input_major = 1, the FP8_2BIT
 * engine's layout); the values do not depend on the layout. */
int fx_synth_channelized(fx_ctx* ctx, uint64_t seed, int nchan, long ntime, int ninp,
                         double rho, double thresh, int input_major, uint8_t* d_codes);
/* 2-bit channelized codes (either layout) -> [c][t][i] FP32 (fx_cf32) or
 * FP16 (half2) samples through the level table: the input of the FP32 and
 * FP16 engines for channelized data. */
int fx_codes_to_chan(fx_ctx* ctx, const uint8_t* d_codes, int input_major, int nchan, long ntime,
                     int ninp, int out_prec, void* d_out);

/* ---- primitives ------------------------------------------------------------ */
/* a1: 2-bit unpack through {-3.3359,-1,+1,+3.3359}; nchan in {1,2,4,8,16};
 * out[c*ntime + t]. */
int fx_unpack2b(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, float* d_out);
/* a3: forward complex FFT, unnormalized, natural order; n = 2^k, 2..16384. */
int fx_fft_c2c(fx_ctx* ctx, const fx_cf32* d_in, fx_cf32* d_out, int n, long batch);
/* a3: real-input FFT, n = 2^k, 4..32768; out has n/2 (or n/2+1 with
 * keep_nyquist) complex bins per transform. */
int fx_fft_r2c(fx_ctx* ctx, const float* d_in, fx_cf32* d_out, int n, long batch,
               int keep_nyquist);
/* acc += a*b (conj=0, genericAddProduct_32fc) or a*conj(b) (conj=1). */
int fx_complex_mac(fx_ctx* ctx, const fx_cf32* d_a, const fx_cf32* d_b, fx_cf32* d_acc,
                   long n, int conj);
int fx_complex_mul(fx_ctx* ctx, const fx_cf32* d_a, const fx_cf32* d_b, fx_cf32* d_out, long n);
int fx_complex_add(fx_ctx* ctx, const fx_cf32* d_a, const fx_cf32* d_b, fx_cf32* d_out, long n);
/* device copy through a plain load/store kernel (16-byte multiples, aligned);
 * used to calibrate HBM traffic counters */
int fx_copy_kernel(fx_ctx* ctx, const void* d_src, void* d_dst, size_t bytes);
/* (cos, sin) of arg (radians, or turns when normalized != 0). */
int fx_sincos(fx_ctx* ctx, const float* d_arg, fx_cf32* d_out, long n, int normalized);
/* oscillator output (cos, sin)(start + i*inc), phase accumulated in FP64. */
int fx_sincos_generator(fx_ctx* ctx, double start, double inc, fx_cf32* d_out, long n);

/* ---- a8: spectrometer ------------------------------------------------------ */
/* nchan interleaved 2-bit channels of ntime samples; n-point real FFT per
 * channel, window, |X|^2 summed over integrate_ffts FFTs and scaled by
 * 1/integrate_ffts.  d_out[s][c][n/2], s < ntime/(n*integrate_ffts). */
int fx_spectrometer(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                    int integrate_ffts, int window, float* d_out);
/* Overlapped FFTs: hop = n/overlap samples between FFT starts.  continuous=1
 * runs one unbroken sequence (the SPU kernel's convention: spectrum s uses
 * FFTs s*K .. s*K+K-1); continuous=0 starts each spectrum after the last
 * sample of the previous one (the IPP core's convention).  keep_nyquist=1
 * writes n/2+1 bins per spectrum (fft_ssb_points, Settings.h:127), else n/2.
 * *nspec_out (may be NULL) receives the number of spectra written. */
int fx_spectrometer_ex(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                       int overlap, int integrate_ffts, int window, int continuous, int keep_nyquist,
                       float* d_out, long* nspec_out);
/* No counterpart in the reference: fx_spectrometer_ex for the callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Same arguments, same kernels and launch shape, identical spectra;
 * it calls no hipMalloc/hipFree: the window table and the partial spectra
 * live in one static device array of FX_SPEC_NOALLOC_BYTES per GPU (held until
 * the process ends).  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU race.  A shape whose partial spectra need more than
 * FX_SPEC_NOALLOC_BYTES - 128 KiB (nspec x nchan x groups x (8192 or 16384 +
 * 1) floats) is refused with FX_ERR_ARG; the 16-channel, 32768-point
 * whole-stream spectrum needs 34.6 MB. */
#define FX_SPEC_NOALLOC_BYTES (40u << 20)
int fx_spectrometer_ex_noalloc(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                               int overlap, int integrate_ffts, int window, int continuous, int keep_nyquist,
                               float* d_out, long* nspec_out);
/* Cross spectrum of channel chan_a of stream a and channel chan_b of stream b
 * (TaskCoreIPP.cpp:456-487, the cross-polarization product of two sources):
 * the same windowed, overlapped FFT frames and spectrum boundaries as
 * fx_spectrometer_ex (ntime samples per channel in both streams), per
 * spectrum n/2+1 complex bins of (1/integrate_ffts) sum X_a conj(X_b) with
 * FP32 accumulation (FFT order within groups of FFTs, then the groups in
 * order: deterministic); DC and Nyquist hold the product of the two
 * real parts with a zero imaginary part, as the IPP core keeps them.  d_out:
 * nspec x (n/2+1) fx_cf32.  One pass holds at most 1 GiB of frames and
 * transforms of whole spectra (integrate_ffts <= 65535). */
int fx_spectrometer_xpol(fx_ctx* ctx, const uint32_t* d_words_a, int nchan_a, int chan_a,
                         const uint32_t* d_words_b, int nchan_b, int chan_b, size_t ntime, int n, int overlap,
                         int integrate_ffts, int window, int continuous, fx_cf32* d_out, long* nspec_out);

/* ---- a5: phase-cal ------------------------------------------------------- */
typedef struct fx_pcal fx_pcal;
/* PCal::getNew semantics: picks the Trivial, ImplicitShift or Shifting
 * extractor from (bandwidth, spacing, offset). */
int fx_pcal_create(fx_ctx* ctx, double bandwidth_hz, double spacing_hz, int offset_hz,
                   size_t sampleoffset, fx_pcal** pc);
int fx_pcal_destroy(fx_pcal* pc);
int fx_pcal_clear(fx_pcal* pc);
int fx_pcal_adjust_sample_offset(fx_pcal* pc, size_t sampleoffset);
/* fold nsamp 2-bit samples (1-channel packing) starting at sample 0 of d_words */
int fx_pcal_extract_2bit(fx_pcal* pc, const uint32_t* d_words, size_t nsamp);
/* host output, ntones complex per input (ninp x ntones for a multi-input
 * extractor); returns samples integrated via *nsamples */
int fx_pcal_final(fx_pcal* pc, fx_cf64* tones, uint64_t* nsamples);
int fx_pcal_info(fx_pcal* pc, int* kind, int* nbins, int* ntones);
/* graph-capturable split of fx_pcal_final: enqueue the tone extraction into
 * device memory, then (after the stream or graph ran) copy the tones out */
int fx_pcal_final_async(fx_pcal* pc);
int fx_pcal_tones(fx_pcal* pc, fx_cf64* tones, uint64_t* nsamples);
/* ninp inputs with one shared tone configuration and sample position: the
 * fold and the tone extraction are one launch each for all inputs.  Input i
 * starts at d_words + i*in_stride_words; fx_pcal_tones_input reads input i.
 * fx_pcal_create(...) is fx_pcal_create_inputs(..., 1, ...). */
int fx_pcal_create_inputs(fx_ctx* ctx, double bandwidth_hz, double spacing_hz, int offset_hz,
                          size_t sampleoffset, int ninp, fx_pcal** pc);
int fx_pcal_extract_2bit_inputs(fx_pcal* pc, const uint32_t* d_words, size_t in_stride_words,
                                size_t nsamp);
int fx_pcal_tones_input(fx_pcal* pc, int input, fx_cf64* tones, uint64_t* nsamples);

/* ---- real-file front end: VDIF / Mark5B frames -------------------------- */
/* VDIF header fields (format_vdif.c:135-206,2199-2240; VDIF specification):
 * word0 bits 0-29 seconds from the reference epoch, bit 30 legacy (16-byte
 * header), bit 31 invalid; word1 bits 0-23 frame number within the second,
 * bits 24-29 reference epoch (half-years since 2000); word2 bits 0-23 frame
 * length in 8-byte units, bits 24-28 log2(channels), bits 29-31 version;
 * word3 bits 0-15 station, bits 16-25 thread, bits 26-30 bits per sample - 1,
 * bit 31 complex. */
typedef struct {
    uint32_t seconds;
    int legacy, invalid;
    uint32_t frame;
    int epoch;
    int framebytes, headerbytes;
    uint32_t nchan;     /* 2^(log2 channels), up to 2^31 */
    int version;
    int station, thread, nbits, is_complex;
    int mjd;            /* epoch start + seconds (format_vdif.c:143-183 table) */
    int second_of_day;
} fx_vdif_header;
int fx_vdif_decode(const void* header, fx_vdif_header* h);
/* Mark5B header (mark5_format_mark5b.c:43,129-186): word0 sync 0xABADDEED,
 * word1 bits 0-14 frame number within the second (bit 15 test vector),
 * word2 BCD JJJSSSSS (MJD mod 1000, second of day), word3 BCD fraction of a
 * second in its top four digits (0.1 ms units, rounded up to the 156.25 us
 * frame grid as mark5access "unrounds" it). */
typedef struct {
    int sync_ok, frame, tvg;
    int day;            /* MJD mod 1000 */
    int second_of_day;
    double frac_ns;
} fx_mark5b_header;
int fx_mark5b_decode(const void* header, fx_mark5b_header* h);
/* Place nframes frames of framebytes bytes from d_raw (HBM) into per-thread
 * rows: the payload of a frame of thread th (VDIF word3 thread id; 0 for
 * Mark5B) at time (second s, frame number f) goes to slot
 * (s - sec0) * frames_per_second + f - frame0 of row d_thread_row[th]
 * (map_size entries; an entry outside [0, nrows) or a thread id >= map_size
 * means not wanted), payload_words = (framebytes -
 * payload_offset)/4 words per slot, rows out_stride_words apart.  Seconds are
 * VDIF epoch seconds, or day*86400 + second of day for Mark5B.  Per slot,
 * d_valid[(row*nslots + slot)*2 + {0,1}] gets the valid 32-bit word range
 * [start, end) of the payload, as mark5access blanks it: empty for a VDIF
 * frame with the invalid bit or a Mark5 fill word (0x11223344 twice) at either
 * end of the payload (format_vdif.c:2337-2343, blanker_mark5.c:282-320), and
 * for Mark5B the span outside a fill run at one end (blanker_mark5.c:54-200,
 * one 32 KB zone per frame; empty without the sync word).  The straggler test
 * at the start of a Mark5B payload reads the first 32-bit word of the first
 * non-fill 64-bit word; mark5access reads the word 8 bytes earlier, a fill
 * word, and so starts every such range one word late.  By default the
 * mark5access searches are reproduced as they are, including that test and
 * their inexact bisection (it never probes the index its last step lands on,
 * so a boundary can come up to 36 bytes late); FX_PLACE_EXACT_BLANK bisects
 * to the exact fill boundary of a payload with one fill run and tests the
 * right straggler word.  VDIF 2-bit payload
 * words are copied as they are; Mark5B words get the two bits of each sample
 * swapped (sign/magnitude codes -H,+1,-1,+H, mark5_format_mark5b.c:59-96, to
 * this library's -H,-1,+1,+H).  d_counts (4 x u64, accumulated): placed,
 * placed but empty, slot out of range, unmapped thread.  flags:
 * FX_PLACE_CLEAR first empties every slot's range and zeroes d_counts;
 * FX_PLACE_EXACT_BLANK as above.  A slot written by two frames of one call
 * holds an unspecified mix of them.  Asynchronous on the context stream. */
enum { FX_PLACE_CLEAR = 1, FX_PLACE_EXACT_BLANK = 2 };
int fx_frames_place(fx_ctx* ctx, int format, const void* d_raw, size_t nframes, int framebytes,
                    int payload_offset, const int* d_thread_row, int map_size, int nrows, long sec0,
                    long frame0, int frames_per_second, long nslots, uint32_t* d_out, size_t out_stride_words,
                    uint32_t* d_valid, uint64_t* d_counts, int flags);
/* Double-buffered file ingest into HBM (the TaskDispatcher prefill and double
 * buffering, TaskDispatcher.cpp:116-137): a reader thread fills nbuf pinned
 * host buffers of chunk_bytes with pread and copies them to device buffers on
 * the ingest's own stream.  fx_ingest_next makes the context stream wait for
 * the next chunk's copy and returns its device pointer and size (0 at end of
 * file); the chunk stays valid until the next fx_ingest_next call, which
 * also lets its buffer be refilled once the work the caller queued on the
 * context stream meanwhile has finished.  offset skips bytes at the start. */
typedef struct fx_ingest fx_ingest;
int fx_ingest_open(fx_ctx* ctx, const char* path, size_t offset, size_t chunk_bytes, int nbuf,
                   fx_ingest** out);
int fx_ingest_next(fx_ingest* g, const void** d_chunk, size_t* bytes);
int fx_ingest_close(fx_ingest* g);

/* ---- visibility products downstream of the dump ------------------------- */
/* Lag spectra of nrows visibility rows of nchan channels (zerocorr.c:391,
 * 633-645: FFTW_BACKWARD over the channels, unnormalized, then the 1/n_ffts
 * scale): lags[r][m] = scale * sum_k vis[r][k] e^{+2 pi i k m / nchan},
 * natural order (lag j < 0 at m = j + nchan); nchan = 2^k, 2..16384.  The
 * dump's [p][c] layout is nrows = P rows of nchan = C channels. */
int fx_lags(fx_ctx* ctx, const fx_cf32* d_vis, long nrows, int nchan, float scale, fx_cf32* d_lags);
/* Amplitude sqrt(x^2 + y^2) and phase atan2(y, x) of (x, y) = scale * z for
 * n complex values (zerocorr.c:628-630,639-643), computed in FP64; either
 * output may be NULL. */
int fx_amp_phase(fx_ctx* ctx, const fx_cf32* d_z, long n, float scale, float* d_amp, float* d_phase);
/* Phase-cal delay (ns) of each of ninp inputs from its ntones tones
 * (fx_pcal_final layout, ninp x ntones complex FP64) at frequencies
 * tone0_hz + t * tone_step_hz: the m5pcal fit (m5pcal.c:94-150) -- mean of
 * the adjacent-tone phase slopes of the tones within band_valid_hz of
 * band_center_hz, each folded once by 2 pi toward the slope of the two middle
 * tones, wrapped into +-500 / (tone spacing in MHz) ns.  Phases are
 * atan2(im, re).  0 for fewer than two usable tones. */
int fx_pcal_delay(fx_ctx* ctx, const fx_cf64* d_tones, int ninp, int ntones, double tone0_hz,
                  double tone_step_hz, double band_center_hz, double band_valid_hz, double* d_delay_ns);

/* ---- a3/a4/a6/a7: correlator -------------------------------------------- */
typedef struct {
    int    ninp;          /* inputs = stations x polarizations                  */
    int    nfft;          /* real FFT length N; C = N/2 channels               */
    long   chunk_spectra; /* spectra per F->X chunk (channelized buffer depth)  */
    int    prec;          /* FX_PREC_*                                          */
    int    use_phase;     /* apply the FP64 phase model                         */
} fx_corr_config;
typedef struct fx_corr fx_corr;
int fx_corr_create(fx_ctx* ctx, const fx_corr_config* cfg, fx_corr** corr);
int fx_corr_destroy(fx_corr* corr);
/* host array of ninp phase models (copied) */
int fx_corr_set_phase_model(fx_corr* corr, const fx_phase_model* pm);
/* optional per-(channel) multiplier table [ninp][C] (device), applied after
 * the FFT as genericMul_32fc; NULL disables */
int fx_corr_set_multipliers(fx_corr* corr, const fx_cf32* d_mult);
/* host arrays: per-channel [C] and per-input [ninp] power scales applied
 * before FP16/FP8 conversion and removed in the dump (NULL = defaults) */
int fx_corr_set_scales(fx_corr* corr, const float* chan_scale, const float* inp_scale);
/* F + X over spectra [t_first, t_first+nspectra): d_words points at input 0,
 * input i at d_words + i*in_stride_words; spectrum t starts at sample t*N. */
int fx_corr_process(fx_corr* corr, const uint32_t* d_words, size_t in_stride_words,
                    long t_first, long nspectra);
/* X only over channelized data in the engine layout:
 *   FX_PREC_F32 / F16: [c][t][i] complex (fx_cf32 / half2);
 *   FX_PREC_FP8:       [c][i][t] e4m3 pairs (byte re | byte im << 8), scaled
 *                      by chan_scale[c] * inp_scale[i] (removed in the dump),
 *                      nspectra a multiple of 8;
 *   FX_PREC_FP8_2BIT:  [c][i][t] 2-bit codes (byte re | im << 2), nspectra a
 *                      multiple of 16; exact: X = S + (H-1) T with S the signs
 *                      and T the signs of the |x| = H samples, XX^H =
 *                      (2-H) SS^H + (H-1) UU^H + (H-1)(H-2) TT^H, U = S + T,
 *                      three FP8 GEMMs on values in {0, +-1, +-2} with exact
 *                      FP32 sums (up to 2^21 spectra per dump) combined
 *                      exactly by the dump (64-bit integer coefficients) and
 *                      rounded once to FP32 -- the exact sums to nearest-even;
 *   FX_PREC_FP4_2BIT:  the same input and decomposition with the operands in
 *                      FP4-e2m1 (0, +-1, +-2 are e2m1 values): the same
 *                      integer sums, the same visibilities. */
int fx_corr_xengine(fx_corr* corr, const void* d_chan, long nspectra);
/* write visibilities [p][c] complex64, p = i*I - i(i-1)/2 + (j-i), i <= j */
int fx_corr_dump(fx_corr* corr, fx_cf32* d_vis);
int fx_corr_reset(fx_corr* corr);
/* F-stage alone into a caller buffer in the engine layout of out_prec
 * (FP32/F16 [c][t][i], FP8 [c][i][t]) */
int fx_corr_fstage(fx_corr* corr, const uint32_t* d_words, size_t in_stride_words,
                   long t_first, long nspectra, void* d_chan, int out_prec);
/* For real FFT lengths 4096, 8192 and 16384 with an input count (FP32/FP16)
 * or spectrum count (FP8) that is a multiple of 32768/n, fx_corr_fstage is one
 * launch of a transform kernel that writes the engine layout itself (a
 * workgroup transforms 32768/n neighbouring columns and stores one channel of
 * all of them per lane; FX_FSTAGE_DIRECT=0 in the environment at
 * fx_corr_create turns this off).  Otherwise it works in sub-chunks of
 * fx_corr_fstage_depth() spectra (about 64 MB of rows, so the corner turn
 * reads them back from the Infinity Cache): per sub-chunk, the transform
 * kernel writes natural rows into the correlator's row buffer and the corner
 * turn moves them into d_chan.  fx_corr_fstage_launches says which: the number
 * of kernel launches fx_corr_fstage makes for nspectra spectra.  The two
 * sub-chunk launches are exposed one at a time for per-kernel timing: rows of
 * spectra [t_first, t_first + nspectra), nspectra <= depth, then their corner
 * turn to spectra [s0, s0 + nspectra) of a chunk of chunk_spectra spectra in
 * d_chan. */
int fx_corr_fstage_launches(fx_corr* corr, long nspectra, int out_prec);
/* name of the F transform kernel fx_corr_fstage runs on nspectra spectra
 * (k_fstage_direct_p: the persistent direct kernel, launched after a clear of
 * its eight block counters) */
int fx_corr_fstage_name(fx_corr* corr, long nspectra, int out_prec, char* buf, int len);
long fx_corr_fstage_depth(fx_corr* corr);
int fx_corr_fstage_rows(fx_corr* corr, const uint32_t* d_words, size_t in_stride_words,
                        long t_first, long nspectra, int out_prec);
int fx_corr_cornerturn(fx_corr* corr, long nspectra, long s0, long chunk_spectra, int out_prec,
                       void* d_chan);
int fx_corr_info(fx_corr* corr, long* chunk_spectra, size_t* chan_bytes, size_t* acc_bytes);
/* tuning knobs (0 keeps the current value): F-stage sub-chunk depth in
 * spectra (multiple of 8) and the X-engine input tile, 128 or 64 (64 only for
 * the FP32/FP16 engines; smaller tiles waste less MFMA work on small input
 * counts).  Changing the tile clears the accumulators. */
int fx_corr_tune(fx_corr* corr, long sub_spectra, int xtile);
/* X-engine choice of the FP32/FP16 correlators: FX_XENGINE_AUTO (default;
 * FX_XENGINE_RING=0 in the environment at fx_corr_create makes the default
 * FX_XENGINE_PREFETCH) runs the LDS-ring engines where they apply -- exactly
 * 128 inputs with a multiple of KT spectra per call (32 FP16, 16 FP32), or a
 * multiple of 128 inputs with a multiple of KT/2 -- and the register-prefetch
 * engine otherwise; FX_XENGINE_PREFETCH runs the register-prefetch engine for
 * every shape.  FX_XENGINE_LOCKSTEP (or FX_XENGINE_LOCK=1 in the environment
 * at fx_corr_create) runs the ring engines as AUTO does, the large-input one
 * -- a multiple of 128 inputs (2 to 64 tiles per channel), a multiple of
 * 4 KT/2 spectra -- in lockstep groups: the workgroups of one channel's tiles
 * on one XCD, kept within a few stages of each other, so the input blocks
 * they share are fetched from HBM once instead of about three times (slower
 * on MI355X all the same: the engine is not bound by those fetches).  All give
 * bit-identical sums. */
enum { FX_XENGINE_AUTO = 0, FX_XENGINE_PREFETCH = 1, FX_XENGINE_LOCKSTEP = 2 };
int fx_corr_set_xengine(fx_corr* corr, int kind);
/* Form of the direct F kernel (FFT lengths 4096-16384): FX_FSTAGE_AUTO
 * (default) runs its persistent form, one workgroup per CU walking its XCD's
 * blocks from a counter, with FP16 blocks in pairs where the inputs are a
 * multiple of 2F (32-byte store runs); FX_FSTAGE_PERSISTENT the persistent
 * form with single blocks; FX_FSTAGE_BLOCKS one block per workgroup.  The
 * same values in every form (the environment variables FX_FSTAGE_PERSIST=0
 * and FX_FSTAGE_PAIR=0 set the default at creation). */
enum { FX_FSTAGE_AUTO = 0, FX_FSTAGE_PERSISTENT = 1, FX_FSTAGE_BLOCKS = 2 };
int fx_corr_set_fstage(fx_corr* corr, int kind);
/* name of the X-engine kernel fx_corr_xengine / fx_corr_process runs on
 * nspectra spectra of d_chan (NULL: the correlator's own channel buffer) */
int fx_corr_xengine_name(fx_corr* corr, const void* d_chan, long nspectra, char* buf, int len);

/* ---- batched subbands (uneven FFT lengths and input counts) -------------- */
typedef struct {
    int nfft;                  /* N, 256..16384                                    */
    int ninp;                  /* inputs (stations x polarizations)                */
    long nspectra;             /* spectra integrated                               */
    const uint32_t* d_words;   /* input 0; input i at d_words + i*stride_words     */
    size_t stride_words;
    const fx_phase_model* pm;  /* host array of ninp models, or NULL (no rotation) */
    int pcal;                  /* extract phase-cal tones of every input (FP64)    */
    double bandwidth_hz, pcal_spacing_hz;
    int pcal_offset_hz;
} fx_subband_desc;
typedef struct fx_batch fx_batch;
/* Modes (use_graph):
 *  FX_BATCH_EAGER   one correlator per subband, issued one launch at a time;
 *  FX_BATCH_GRAPH   the same per-subband work (reset, F-stage, X-engine, dump,
 *                   phase-cal fold and tone extraction) captured once into a
 *                   hipGraph with one concurrent branch per subband;
 *  FX_BATCH_PERSISTENT  (FP16) one persistent launch, in a hipGraph, walks a
 *                   work table of every subband: phase-cal folds, F groups
 *                   (unpack -> FFT -> rotate -> FP16, stored in the X-engine
 *                   layout [c][t][i] over the whole integration, or as
 *                   natural rows that corner-turn items re-tile), X items
 *                   (channels of a 64 x 64 input tile over the integration or
 *                   one time slice of it), slice reductions, tone
 *                   extractions; an item waits in the kernel for the items
 *                   whose output it reads (they come earlier in the table), a
 *                   bounded wait that fails the run (fx_batch_sync) if it ever
 *                   expires;
 *  FX_BATCH_PHASED  (FP16) the same table as four launches (folds and F
 *                   groups, corner turns, X items, reductions and tones), no
 *                   waits.
 * The persistent modes keep each subband's channelized samples for its whole
 * integration in HBM (nfft/2 x nspectra x ninp x 4 bytes, and as much again
 * for natural rows) and fold phase-cal of up to 4096 bins (multiples of 16).
 * fx_batch_run replays the graph (or issues the eager launches) with one call.
 * tune_file: autotuner cache (NULL: defaults); entries missing from it are
 * tuned on the spot and appended -- per-subband F sub-chunk depth and X tile
 * for the first two modes (fx_autotune); natural rows or direct stores, F
 * groups per item, channels per X item and time slices for the persistent
 * modes ("P" lines, timed on the subband's share of the device). */
enum { FX_BATCH_EAGER = 0, FX_BATCH_GRAPH = 1, FX_BATCH_PERSISTENT = 2, FX_BATCH_PHASED = 3 };
int fx_batch_create(fx_ctx* ctx, const fx_subband_desc* sb, int nsub, int prec, int use_graph,
                    const char* tune_file, fx_batch** batch);
int fx_batch_run(fx_batch* batch);
int fx_batch_sync(fx_batch* batch);
/* device visibilities [p][c] of subband s (valid after a run) */
const fx_cf32* fx_batch_vis(fx_batch* batch, int s);
/* host tones of input i of subband s (after fx_batch_sync) */
int fx_batch_pcal(fx_batch* batch, int s, int i, fx_cf64* tones, int* ntones);
/* chosen tuning of subband s */
int fx_batch_tuning(fx_batch* batch, int s, long* sub_spectra, int* xtile, long* chunk_spectra);
/* persistent modes: natural rows + corner turns (1) or direct stores (0),
 * channels per X item, F groups per item, X time slices and X tile width (64,
 * or 128: one tile of 65 to 128 inputs) of subband s; the work items,
 * workgroups and dynamic LDS bytes of the launch */
int fx_batch_persist_tuning(fx_batch* batch, int s, int* natural_rows, int* chan_per_item, int* groups_per_item,
                            int* time_slices, int* xtile);
int fx_batch_persist_info(fx_batch* batch, int* nitems, int* nblocks, size_t* lds_bytes);
int fx_batch_destroy(fx_batch* batch);

/* Autotuner: times the F-stage + X-engine of one chunk of chunk_spectra
 * spectra of an (nfft, ninp, prec) subband on synthetic input for every
 * candidate (sub-chunk depth x X tile) and appends the fastest to tune_file
 * ("nfft ninp prec chunk sub xtile ms" lines, looked up by the first four). */
int fx_autotune(fx_ctx* ctx, int nfft, int ninp, int prec, long chunk_spectra, const char* tune_file,
                long* sub_spectra, int* xtile);

/* ---- diagnostics --------------------------------------------------------- */
/* MFMA lane-layout probe: which = 0 (f32 32x32x2), 1 (f16 32x32x16),
 * 2 (fp8 32x32x64); writes 32x32 floats D[i][j] for A[i][k] = i+1 when
 * k == 0 else 0, B[k][j] = j+1 when k == 0 else 0 (so D[i][j] = (i+1)(j+1)). */
int fx_mfma_probe(fx_ctx* ctx, int which, float* host_out);

#ifdef __cplusplus
}
#endif
#endif

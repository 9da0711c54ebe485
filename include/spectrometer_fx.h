/*
 * spectrometer_fx.h -- the software spectrometer's compute seams.
 *
 * 1. SPU control block (cell_spectrometer.h:104-115): spectrometer_cb_run()
 *    performs what one mailbox wake-up of spectrometer_spu.c:97-224 does --
 *    unpack, sin^2 window, N-point FFT, power accumulation over integrate_ffts
 *    overlapped FFTs, one spectrum written to spectrum_dst per integration --
 *    and returns the three mailbox words (t_spu, #spectra, #FFTs).
 *
 * 2. TaskCore (TaskCore.h:36-89): fx_taskcore_* is the asynchronous
 *    prepare/run/join/finalize/combineResults core a TaskDispatcher drives;
 *    run() returns at once and the work proceeds on the core's own HIP
 *    stream, join() blocks and returns the number of complete spectra.
 *
 * Conventions fixed here (deviations from the SPU code are its defects):
 * raw input is 2-bit offset-binary samples in the VDIF layout
 * (samples_per_rawbyte = 4; spectrometer_spu.c:329-343 decodes bytes as
 * 8-bit whatever the setting), power is (1/K) sum |X|^2 with the 1/K applied
 * after squaring (TaskCoreIPP.cpp:124,493; spectrometer_spu.c:137,269 squares
 * the pre-scaled value), and the twiddles are complete (spectrometer_spu.c:
 * 287-290 leaves W[0].im unset).
 */
#ifndef SPECTROMETER_FX_H
#define SPECTROMETER_FX_H

#include <stddef.h>
#include <stdint.h>

#include "cellspe_fx.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef FX_ADDR64_DEFINED
#define FX_ADDR64_DEFINED
typedef union {
    unsigned long long ull;
    unsigned int ui[2];
    void* p;
} addr64;
#endif

/* cell_spectrometer.h:104-115 */
typedef struct control_block_tt {
    unsigned int rank __attribute__((aligned(128)));
    unsigned int terminate_flag;      /* nonzero: spectrometer_cb_run returns 0 at once   */
    unsigned int fft_points;          /* N, power of two, 256..32768                       */
    unsigned int fft_points_log2;
    unsigned int overlap_factor;      /* 2 for 50 %                                        */
    unsigned int integrate_ffts;      /* FFTs summed per spectrum                          */
    unsigned int samples_per_rawbyte; /* 4: 2-bit samples                                  */
    addr64 rawif_src;                 /* host raw bytes                                    */
    unsigned int rawif_bytecount;
    addr64 spectrum_dst;              /* host: per spectrum N (re=power, im=0) float pairs */
} control_block_t __attribute__((aligned(128)));

/* the three outbound mailbox words (spectrometer_spu.c:217-220) */
typedef struct {
    unsigned int t_spu;          /* device time in CELLDSP_TIMEBASE ticks (10 ns) */
    unsigned int num_spectra;
    unsigned int total_ffts;
} spectrometer_mbox_t;

/* One control-block run.  spectrum_dst receives num_spectra blocks of
 * fft_points (power, 0) float pairs: bins 0..N/2 computed (Nyquist included),
 * bins N/2+1..N-1 the mirror |X(N-k)|^2 of a complex transform of real data.
 * Returns 0, or -1 on invalid settings / device errors (fx_last_error()). */
int spectrometer_cb_run(const control_block_t* cb, spectrometer_mbox_t* mbox);

/* ---- TaskCore ------------------------------------------------------------ */
/* the swspect_settings_t fields the compute core reads (Settings.h:59-159) */
typedef struct {
    int fft_points;
    int fft_overlap_factor;
    int core_overlapped_ffts;   /* FFTs per spectrum (averaged_overlapped_ffts)       */
    int window;                 /* FX_WIN_*                                           */
    int bits_per_sample;        /* 2                                                  */
    int source_channels;        /* interleaved channels per raw word group: 1..16     */
    int use_channel;            /* 1-based channel to process; 0 = all channels       */
    int extract_pcal;           /* phase-cal per spectrum (single-channel sources)    */
    double samplingfreq;        /* Hz */
    double pcaloffsethz;
    double pcalharmonicshz;
    int use_channel_file2;      /* 1-based channel of sources 2.. in fx_taskcore_run_buffers
                                   (use_channel_file2, Settings.h:68); 0: as use_channel */
} fx_spectrometer_settings;

typedef struct fx_taskcore fx_taskcore;
/* TaskCore::prepare.  Phase-cal parameters as swspectrometer.cpp:178-196:
 * tonebins = fs/harmonics, rotatorlen = fs/gcd(fs, offset) (tonebins when the
 * offset is 0); extraction is switched off (pcal length 0) when rotatorlen
 * does not divide fft_points, and refused when tonebins does not divide
 * rotatorlen (the IPP core would drop the remainder samples). */
int fx_taskcore_create(int rank, const fx_spectrometer_settings* settings, fx_taskcore** core);
/* TaskCore::run: processes rawbytes of raw data asynchronously.  Spectrum s
 * integrates core_overlapped_ffts FFTs hopping fft_points/overlap samples and
 * the next spectrum starts after its last sample (TaskCoreIPP.cpp:495-510).
 * spectra_out receives per spectrum and output channel fft_points/2+1 floats
 * (fft_ssb_points, mean power).  pcal_out (may be NULL) receives per spectrum
 * tonebins complex values: the rotated, folded phase-cal pulse of the
 * spectrum's non-overlapped FFT inputs (TaskCoreIPP.cpp:413-418,633-693),
 * accumulated in FP64.  Buffers stay the caller's and must remain valid until
 * join().  Returns 0 when the work was started. */
int fx_taskcore_run(fx_taskcore* core, const void* raw, size_t rawbytes, float* spectra_out,
                    size_t spectra_capacity_floats, fx_cf64* pcal_out, size_t pcal_capacity);
/* Buffer (Buffer.h): a caller-owned byte array; the core reads `length`
 * bytes of an input and sets `length` of an output (Buffer::setLength) */
typedef struct {
    void* data;        /* Buffer::getData()      */
    size_t allocated;  /* Buffer::getAllocated() */
    size_t length;     /* Buffer::getLength()    */
} fx_buffer;
/* TaskCore::run(Buffer** inbuf, Buffer** outbuf, Buffer** xpolbuf,
 * Buffer** pcalbuf) (TaskCore.h:58), the IPP core's doMaths
 * (TaskCoreIPP.cpp:356-560), asynchronous until join().  nsources raw
 * inputs (1..FX_TC_MAX_SOURCES); source 0 uses channel use_channel (1 if 0)
 * of its source_channels, the others use_channel_file2.  The sources are
 * processed while every one of them has samples left: spectra of
 * core_overlapped_ffts FFTs hopping fft_points/overlap samples, the next
 * spectrum after the last sample of the previous one.  Per spectrum:
 *   outbuf[s]  fft_points/2+1 float32, the mean power (1/K) sum |X|^2;
 *   xpolbuf[x] (NULL: none) fft_points/2+1 complex float32, (1/K) sum
 *              X_a conj(X_b) for pair x of (a, b), a < b, in the order
 *              (0,1), (0,2), ..., (1,2), ...; DC and Nyquist the product of
 *              the real parts, imaginary part 0;
 *   pcalbuf[s] (NULL: none; single-channel sources) tonebins complex
 *              float32, the folded phase-cal pulse of the spectrum's
 *              non-overlapped FFT inputs (FFT index a multiple of the
 *              overlap factor), accumulated in FP64.
 * When the shortest source ends with 1..K-1 whole FFTs past the last full
 * spectrum, those are written as one more, partial spectrum of unscaled sums
 * (TaskCoreIPP.cpp:519-524).  Each output's length is set to its spectra's
 * bytes after the work; allocated bytes that cannot hold them fail the run
 * (join() returns -1).  The buffers stay the caller's until join(). */
#define FX_TC_MAX_SOURCES 8
int fx_taskcore_run_buffers(fx_taskcore* core, int nsources, fx_buffer* const* inbuf, fx_buffer* const* outbuf,
                            fx_buffer* const* xpolbuf, fx_buffer* const* pcalbuf);
/* TaskCore::combineResults(Buffer** outbuf, Buffer** pcalbuf) after
 * fx_taskcore_run_buffers (TaskCoreIPP.cpp:311-345): outbuf[s] += the first
 * spectrum of source s, outbuf[nsources + x] += the first cross spectrum of
 * pair x (complex float32), pcalbuf[s] += the first phase-cal (may be NULL);
 * each length set to one spectrum's bytes. */
int fx_taskcore_combine_buffers(fx_taskcore* core, fx_buffer* const* outbuf, fx_buffer* const* pcalbuf);
/* TaskCore::join: -1 on failure, else the number of spectra written
 * (fx_taskcore_run: complete spectra; fx_taskcore_run_buffers: including a
 * trailing partial spectrum) */
int fx_taskcore_join(fx_taskcore* core);
/* TaskCore::combineResults (TaskCoreIPP.cpp:311-345): out[k] += this core's
 * last spectra, pcal[k] += its last phase-cal (pcal may be NULL) */
int fx_taskcore_combine(fx_taskcore* core, float* out, size_t nfloats, fx_cf64* pcal, size_t npcal);
/* complex values per spectrum of phase-cal output (0 if not extracting) */
int fx_taskcore_pcal_length(fx_taskcore* core);
/* TaskCore::finalize */
int fx_taskcore_finalize(fx_taskcore* core);
int fx_taskcore_destroy(fx_taskcore* core);

/* ---- the spectrometer application (TaskDispatcher + sinks) --------------- */
/* Settings of one run (swspect_settings_t, Settings.h:59-159, and the INI
 * keys UseFile1Channel/UseFile2Channel, swspectrometer.cpp:140-160). */
typedef struct {
    fx_spectrometer_settings core;  /* use_channel is ignored: see use_channel_file */
    int format;                     /* FX_FMT_RAW, FX_FMT_VDIF (one thread per file) or FX_FMT_MARK5B;
                                       frames in time order without gaps */
    int framebytes, headerbytes;    /* frame and header sizes (Mark5B: 10016 and 16) */
    int frames_per_second;          /* frames per second of the file */
    int use_channel_file[2];        /* 1-based channel of input 1 and input 2 */
    int calc_xpol;                  /* cross spectrum of the two selected channels */
    int spectra_per_chunk;          /* whole spectra per double-buffered input chunk */
    const char* basefilename[2];    /* sink names: <base>_swspec.bin, <base>_pcal.bin,
                                       <base 1>_xpol_swspec.bin; with two inputs [1] must be
                                       given and differ from [0] (else FX_ERR_ARG) */
} fx_swspec_settings;

typedef struct {
    long spectra;                   /* spectra written per sink */
    long chunks;                    /* input chunks read */
    double seconds;                 /* wall time of the processing loop */
    double realtime_factor;         /* signal time of the spectra / seconds: spectra x
                                       ((integrate_ffts-1) x fft_points/overlap + fft_points) / fs / seconds */
} fx_swspec_stats;

/* The TaskDispatcher loop (TaskDispatcher.cpp:38-230) over one or two input
 * files (input2 may be NULL): the files are streamed into HBM through the
 * double-buffered ingest (fx_ingest_*), VDIF frames placed by
 * fx_frames_place, and every whole spectrum -- TaskCore spectra of
 * (integrate_ffts - 1) * fft_points/overlap + fft_points samples, counted
 * from the start of the file and independent of the chunking -- written to
 * the binary sinks: fft_points/2+1 float32 per spectrum of each input's
 * selected channel, the cross spectrum of the two (fft_points/2+1 complex
 * float32, fx_spectrometer_xpol) and the phase-cal (tonebins complex
 * float32).  The run ends when either input ends.  Samples of framed inputs
 * that the blanker marks invalid (fx_frames_place ranges; a missing frame is
 * invalid too) decode as zeros, as mark5access decodes them, in the spectra,
 * the cross spectrum and the phase-cal.  Returns 0 or an FX_ERR_* code
 * (fx_last_error()). */
int fx_swspectrometer_run(const fx_swspec_settings* settings, const char* input1, const char* input2,
                          fx_swspec_stats* stats);

#ifdef __cplusplus
}
#endif
#endif

// fx_api.cpp -- host side of the F-X engine C-ABI (include/cellspe_fx.h).
//
// Owns the device context (stream, FP64-derived twiddle table, window
// tables), argument checking, chunking of the correlator time axis and the
// phase-cal extractor state.  No compute runs on the host: every entry point
// enqueues the gfx950 kernels of fx_fft.hip, fx_xengine.hip and fx_misc.hip.
#include "fx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

void fx_set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char* fx_last_error(void) { return g_err; }

#define FX_CHECK_LAUNCH(rc)                                                     \
    do {                                                                        \
        int _rc = (rc);                                                         \
        if (_rc != FX_OK) {                                                     \
            if (_rc == FX_ERR_HIP)                                              \
                fx_set_error("%s:%d: kernel launch failed: %s", __FILE__,       \
                             __LINE__, hipGetErrorString(hipGetLastError()));   \
            else if (!g_err[0])                                                 \
                fx_set_error("%s:%d: unsupported size", __FILE__, __LINE__);    \
            return _rc;                                                         \
        }                                                                       \
    } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct fx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    float2* tw = nullptr;                       // FX_TW_N entries
    std::map<std::pair<int, int>, float*> windows;
    float* scratch = nullptr;                   // spectrometer partials
    size_t scratch_bytes = 0;
    std::mutex mu;
};

static int ctx_device(fx_ctx* ctx)
{
    FX_REQUIRE(ctx != nullptr, FX_ERR_ARG, "null context");
    FX_HIP_TRY(hipSetDevice(ctx->device));
    return FX_OK;
}

extern "C" int fx_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int fx_device_name(int device, char* buf, int len)
{
    hipDeviceProp_t p;
    FX_HIP_TRY(hipGetDeviceProperties(&p, device));
    snprintf(buf, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    return FX_OK;
}

extern "C" int fx_ctx_create(int device, fx_ctx** out)
{
    FX_REQUIRE(out != nullptr, FX_ERR_ARG, "fx_ctx_create: null output");
    *out = nullptr;
    int ndev = 0;
    FX_HIP_TRY(hipGetDeviceCount(&ndev));
    FX_REQUIRE(device >= 0 && device < ndev, FX_ERR_STATE,
               "fx_ctx_create: device %d not present (%d HIP devices)", device, ndev);
    FX_HIP_TRY(hipSetDevice(device));
    fx_ctx* c = new fx_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        fx_set_error("fx_ctx_create: stream creation failed");
        return FX_ERR_HIP;
    }
    // twiddle table derived in FP64, rounded once to FP32
    std::vector<float2> h(FX_TW_N);
    for (int e = 0; e < FX_TW_N; e++) {
        double a = -2.0 * M_PI * (double)e / (double)FX_TW_N;
        h[e] = make_float2((float)cos(a), (float)sin(a));
    }
    if (hipMalloc(&c->tw, sizeof(float2) * FX_TW_N) != hipSuccess ||
        hipMemcpy(c->tw, h.data(), sizeof(float2) * FX_TW_N, hipMemcpyHostToDevice) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        fx_set_error("fx_ctx_create: twiddle table upload failed");
        return FX_ERR_NOMEM;
    }
    *out = c;
    return FX_OK;
}

extern "C" int fx_ctx_destroy(fx_ctx* ctx)
{
    if (!ctx) return FX_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->windows) hipFree(kv.second);
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->tw) hipFree(ctx->tw);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return FX_OK;
}

extern "C" int fx_ctx_sync(fx_ctx* ctx)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FX_OK;
}

extern "C" void* fx_ctx_stream(fx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

const float2* fx_ctx_twiddles(fx_ctx* ctx) { return ctx->tw; }
hipStream_t fx_ctx_hip_stream(fx_ctx* ctx) { return ctx->stream; }
int fx_ctx_device_of(fx_ctx* ctx) { return ctx->device; }

// ---------------------------------------------------------------------------
// memory / events
// ---------------------------------------------------------------------------
extern "C" int fx_malloc(fx_ctx* ctx, size_t bytes, void** dptr)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(dptr != nullptr, FX_ERR_ARG, "fx_malloc: null output");
    *dptr = nullptr;
    if (bytes == 0) return FX_OK;
    if (hipMalloc(dptr, bytes) != hipSuccess) {
        hipGetLastError();
        fx_set_error("fx_malloc: %zu bytes failed", bytes);
        return FX_ERR_NOMEM;
    }
    return FX_OK;
}

extern "C" int fx_free(fx_ctx* ctx, void* dptr)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    if (dptr) FX_HIP_TRY(hipFree(dptr));
    return FX_OK;
}

extern "C" int fx_memcpy_h2d(fx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    if (!bytes) return FX_OK;
    FX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FX_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FX_OK;
}

extern "C" int fx_memcpy_d2h(fx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    if (!bytes) return FX_OK;
    FX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FX_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FX_OK;
}

extern "C" int fx_memcpy_d2d(fx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    if (!bytes) return FX_OK;
    FX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return FX_OK;
}

extern "C" int fx_memset(fx_ctx* ctx, void* dst, int value, size_t bytes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    if (!bytes) return FX_OK;
    // whole aligned words through the fill kernel, so the clear can be
    // captured into a hipGraph (memset nodes are kept out of graphs)
    if ((uintptr_t)dst % 4 == 0 && bytes % 4 == 0) {
        const uint32_t b = (uint32_t)(value & 0xff);
        if (fxk_fill32(ctx->stream, dst, b * 0x01010101u, bytes / 4) != FX_OK) {
            fx_set_error("fx_memset: fill launch failed");
            return FX_ERR_HIP;
        }
        return FX_OK;
    }
    FX_HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return FX_OK;
}

extern "C" int fx_event_create(fx_ctx* ctx, void** ev)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    hipEvent_t e;
    FX_HIP_TRY(hipEventCreate(&e));
    *ev = (void*)e;
    return FX_OK;
}

extern "C" int fx_event_record(fx_ctx* ctx, void* ev)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_HIP_TRY(hipEventRecord((hipEvent_t)ev, ctx->stream));
    return FX_OK;
}

extern "C" int fx_event_elapsed_ms(void* a, void* b, float* ms)
{
    FX_HIP_TRY(hipEventSynchronize((hipEvent_t)b));
    FX_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return FX_OK;
}

extern "C" int fx_event_destroy(void* ev)
{
    if (ev) FX_HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return FX_OK;
}

// ---------------------------------------------------------------------------
// synthetic generators
// ---------------------------------------------------------------------------
extern "C" int fx_synth_station(fx_ctx* ctx, uint64_t seed, int station, double rho, long delay,
                                double tone_amp, double tone_freq_norm, double thresh,
                                size_t nsamp, uint32_t* d_words)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(delay >= 0 && delay <= 64, FX_ERR_ARG, "fx_synth_station: delay %ld not in [0,64]",
               delay);
    FX_CHECK_LAUNCH(fxk_synth_station(ctx->stream, seed, station, rho, delay, tone_amp,
                                      tone_freq_norm, thresh, nsamp, d_words));
    return FX_OK;
}

static bool valid_nchan(int n) { return n == 1 || n == 2 || n == 4 || n == 8 || n == 16; }

extern "C" int fx_synth_multichannel(fx_ctx* ctx, uint64_t seed, int nchan, const double* amp,
                                     const double* freq, double thresh, size_t ntime,
                                     uint32_t* d_words)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(valid_nchan(nchan), FX_ERR_ARG, "fx_synth_multichannel: nchan %d", nchan);
    FX_CHECK_LAUNCH(fxk_synth_multichannel(ctx->stream, seed, nchan, amp, freq, thresh, ntime, d_words));
    return FX_OK;
}

extern "C" int fx_synth_channelized(fx_ctx* ctx, uint64_t seed, int nchan, long ntime, int ninp,
                                    double rho, double thresh, int input_major, uint8_t* d_codes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_synth_channelized(ctx->stream, seed, nchan, ntime, ninp, rho, thresh,
                                          input_major ? 1 : 0, d_codes));
    return FX_OK;
}

extern "C" int fx_codes_to_chan(fx_ctx* ctx, const uint8_t* d_codes, int input_major, int nchan, long ntime,
                                int ninp, int out_prec, void* d_out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(out_prec == FX_PREC_F32 || out_prec == FX_PREC_F16, FX_ERR_ARG,
               "fx_codes_to_chan: output precision %d (FP32 or FP16)", out_prec);
    FX_CHECK_LAUNCH(fxk_codes_to_chan(ctx->stream, d_codes, input_major ? 1 : 0, nchan, ntime, ninp, out_prec,
                                      d_out));
    return FX_OK;
}

// ---------------------------------------------------------------------------
// primitives
// ---------------------------------------------------------------------------
static bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

extern "C" int fx_unpack2b(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, float* d_out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(valid_nchan(nchan), FX_ERR_ARG, "fx_unpack2b: nchan %d not in {1,2,4,8,16}", nchan);
    FX_REQUIRE(nchan != 1 || ((uintptr_t)d_out % 16) == 0, FX_ERR_ARG,
               "fx_unpack2b: output must be 16-byte aligned");
    FX_CHECK_LAUNCH(fxk_unpack2b(ctx->stream, d_words, ntime, nchan, d_out));
    return FX_OK;
}

extern "C" int fx_fft_c2c(fx_ctx* ctx, const fx_cf32* d_in, fx_cf32* d_out, int n, long batch)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(is_pow2(n) && n >= 2 && n <= 16384, FX_ERR_ARG, "fx_fft_c2c: n=%d", n);
    FX_REQUIRE(batch >= 0, FX_ERR_ARG, "fx_fft_c2c: batch=%ld", batch);
    if (batch == 0) return FX_OK;
    FX_CHECK_LAUNCH(fxk_fft_c2c(ctx->stream, ctx->tw, (const float2*)d_in, (float2*)d_out, n, batch));
    return FX_OK;
}

// ---------------------------------------------------------------------------
// visibility products (zerocorr.c:620-645, m5pcal.c:94-150)
// ---------------------------------------------------------------------------
extern "C" int fx_lags(fx_ctx* ctx, const fx_cf32* d_vis, long nrows, int nchan, float scale, fx_cf32* d_lags)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(d_vis && d_lags, FX_ERR_ARG, "fx_lags: null argument");
    FX_REQUIRE(is_pow2(nchan) && nchan >= 2 && nchan <= 16384, FX_ERR_ARG, "fx_lags: nchan=%d", nchan);
    FX_REQUIRE(nrows >= 0, FX_ERR_ARG, "fx_lags: nrows=%ld", nrows);
    if (nrows == 0) return FX_OK;
    FX_CHECK_LAUNCH(fxk_fft_c2c(ctx->stream, ctx->tw, (const float2*)d_vis, (float2*)d_lags, nchan, nrows, 1,
                                scale));
    return FX_OK;
}

extern "C" int fx_amp_phase(fx_ctx* ctx, const fx_cf32* d_z, long n, float scale, float* d_amp, float* d_phase)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(d_z && (d_amp || d_phase) && n >= 0, FX_ERR_ARG, "fx_amp_phase: bad argument");
    FX_CHECK_LAUNCH(fxk_amp_phase(ctx->stream, (const float2*)d_z, n, scale, d_amp, d_phase));
    return FX_OK;
}

extern "C" int fx_pcal_delay(fx_ctx* ctx, const fx_cf64* d_tones, int ninp, int ntones, double tone0_hz,
                             double tone_step_hz, double band_center_hz, double band_valid_hz, double* d_delay_ns)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(d_tones && d_delay_ns && ninp >= 0 && ntones >= 0, FX_ERR_ARG, "fx_pcal_delay: bad argument");
    FX_REQUIRE(tone_step_hz != 0.0 && std::isfinite(tone_step_hz) && std::isfinite(tone0_hz), FX_ERR_ARG,
               "fx_pcal_delay: tone spacing %g Hz", tone_step_hz);
    FX_CHECK_LAUNCH(fxk_pcal_delay(ctx->stream, (const double*)d_tones, ninp, ntones, tone0_hz, tone_step_hz,
                                   band_center_hz, band_valid_hz, d_delay_ns));
    return FX_OK;
}

extern "C" int fx_fft_r2c(fx_ctx* ctx, const float* d_in, fx_cf32* d_out, int n, long batch,
                          int keep_nyquist)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(is_pow2(n) && n >= 4 && n <= FX_MAX_N, FX_ERR_ARG, "fx_fft_r2c: n=%d", n);
    FX_REQUIRE(d_in != (const float*)d_out, FX_ERR_ARG, "fx_fft_r2c: in-place not supported");
    if (batch <= 0) return batch == 0 ? FX_OK : FX_ERR_ARG;
    FX_CHECK_LAUNCH(fxk_fft_r2c(ctx->stream, ctx->tw, d_in, (float2*)d_out, n, batch, keep_nyquist));
    return FX_OK;
}

extern "C" int fx_complex_mac(fx_ctx* ctx, const fx_cf32* a, const fx_cf32* b, fx_cf32* acc,
                              long n, int conj)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_complex_mac(ctx->stream, (const float2*)a, (const float2*)b, (float2*)acc, n, conj));
    return FX_OK;
}

extern "C" int fx_complex_mul(fx_ctx* ctx, const fx_cf32* a, const fx_cf32* b, fx_cf32* out, long n)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_complex_mul(ctx->stream, (const float2*)a, (const float2*)b, (float2*)out, n));
    return FX_OK;
}

extern "C" int fx_complex_add(fx_ctx* ctx, const fx_cf32* a, const fx_cf32* b, fx_cf32* out, long n)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_complex_add(ctx->stream, (const float2*)a, (const float2*)b, (float2*)out, n));
    return FX_OK;
}

extern "C" int fx_copy_kernel(fx_ctx* ctx, const void* src, void* dst, size_t bytes)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0, FX_ERR_ARG,
               "fx_copy_kernel: 16-byte multiples and alignment required");
    FX_CHECK_LAUNCH(fxk_copy(ctx->stream, src, dst, bytes));
    return FX_OK;
}

extern "C" int fx_sincos(fx_ctx* ctx, const float* arg, fx_cf32* out, long n, int normalized)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_sincos(ctx->stream, arg, (float2*)out, n, normalized));
    return FX_OK;
}

extern "C" int fx_sincos_generator(fx_ctx* ctx, double start, double inc, fx_cf32* out, long n)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_sincos_gen(ctx->stream, start, inc, (float2*)out, n));
    return FX_OK;
}

// ---------------------------------------------------------------------------
// spectrometer
// ---------------------------------------------------------------------------
// the window table of (kind, n), computed on the device once per context
static float* get_window(fx_ctx* ctx, int kind, int n)
{
    auto key = std::make_pair(kind, n);
    auto it = ctx->windows.find(key);
    if (it != ctx->windows.end()) return it->second;
    float* d = nullptr;
    if (hipMalloc(&d, sizeof(float) * (size_t)n) != hipSuccess) {
        hipGetLastError();
        return nullptr;
    }
    if (fxk_window(ctx->stream, kind, n, d) != FX_OK) {
        hipFree(d);
        return nullptr;
    }
    ctx->windows[key] = d;
    return d;
}

static int ensure_scratch(fx_ctx* ctx, size_t bytes)
{
    if (ctx->scratch_bytes >= bytes) return FX_OK;
    FX_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) {
        hipGetLastError();
        fx_set_error("scratch allocation of %zu bytes failed", bytes);
        return FX_ERR_NOMEM;
    }
    ctx->scratch_bytes = bytes;
    return FX_OK;
}

extern "C" int fx_spectrometer(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                               int integrate_ffts, int window, float* d_out)
{
    return fx_spectrometer_ex(ctx, d_words, ntime, nchan, n, 1, integrate_ffts, window, 0, 0, d_out, nullptr);
}

int fx_spec_make_plan(size_t ntime, int nchan, int n, int overlap, int integrate_ffts, int window, int continuous,
                      fx_spec_plan* plan)
{
    FX_REQUIRE(valid_nchan(nchan), FX_ERR_ARG, "fx_spectrometer: nchan %d", nchan);
    FX_REQUIRE(is_pow2(n) && n >= 256 && n <= FX_MAX_N, FX_ERR_ARG, "fx_spectrometer: n=%d", n);
    FX_REQUIRE(integrate_ffts >= 1, FX_ERR_ARG, "fx_spectrometer: integrate_ffts=%d", integrate_ffts);
    FX_REQUIRE(overlap >= 1 && n % overlap == 0, FX_ERR_ARG, "fx_spectrometer: overlap %d", overlap);
    FX_REQUIRE(window >= FX_WIN_NONE && window <= FX_WIN_HANN, FX_ERR_ARG,
               "fx_spectrometer: window %d", window);
    const long hop = n / overlap;
    long nspec, spec_stride;
    if (continuous) {
        // one unbroken run of overlapped FFTs (spectrometer_spu.c:171-210)
        long nfft = ntime >= (size_t)n ? (long)((ntime - (size_t)n) / (size_t)hop) + 1 : 0;
        nspec = nfft / integrate_ffts;
        spec_stride = (long)integrate_ffts * hop;
    } else {
        // spectra do not share samples (TaskCoreIPP.cpp:495-510)
        spec_stride = (long)(integrate_ffts - 1) * hop + n;
        nspec = (long)(ntime / (size_t)spec_stride);
    }
    // split each spectrum's FFTs over `groups` workgroups so the grid holds
    // at least ~512 workgroups; each writes one partial spectrum tile
    const int M = n / 2;
    const int FM = (M >= 16384) ? 16384 : 8192;
    const int F = FM / M;
    const long per_spec_units = (integrate_ffts + F - 1) / F;
    const long want = nspec > 0 ? 512 / (nspec * nchan) + 1 : 1;
    plan->nspec = nspec;
    plan->spec_stride = spec_stride;
    plan->hop = hop;
    plan->groups = (int)std::min<long>(std::max<long>(want, 1), per_spec_units);
    plan->part_bytes = (size_t)nspec * nchan * plan->groups * (size_t)(FM + F) * sizeof(float);
    return FX_OK;
}

extern "C" int fx_spectrometer_ex(fx_ctx* ctx, const uint32_t* d_words, size_t ntime, int nchan, int n,
                                  int overlap, int integrate_ffts, int window, int continuous,
                                  int keep_nyquist, float* d_out, long* nspec_out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    fx_spec_plan pl;
    rc = fx_spec_make_plan(ntime, nchan, n, overlap, integrate_ffts, window, continuous, &pl);
    if (rc) return rc;
    if (nspec_out) *nspec_out = pl.nspec;
    if (pl.nspec == 0) return FX_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    float* win = get_window(ctx, window, n);
    FX_REQUIRE(win != nullptr, FX_ERR_NOMEM, "fx_spectrometer: window table");
    rc = ensure_scratch(ctx, pl.part_bytes);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_spectrometer(ctx->stream, ctx->tw, win, d_words, pl.hop, pl.spec_stride, nchan, n,
                                     integrate_ffts, (int)pl.nspec, ctx->scratch, pl.groups, keep_nyquist ? 1 : 0,
                                     d_out));
    return FX_OK;
}

extern "C" int fx_spectrometer_xpol(fx_ctx* ctx, const uint32_t* d_words_a, int nchan_a, int chan_a,
                                    const uint32_t* d_words_b, int nchan_b, int chan_b, size_t ntime, int n,
                                    int overlap, int integrate_ffts, int window, int continuous, fx_cf32* d_out,
                                    long* nspec_out)
{
    return fx_spec_xpol_impl(ctx, d_words_a, nullptr, nchan_a, chan_a, d_words_b, nullptr, nchan_b, chan_b, ntime, n,
                             overlap, integrate_ffts, window, continuous, d_out, nspec_out);
}

int fx_spec_xpol_impl(fx_ctx* ctx, const uint32_t* d_words_a, const uint8_t* d_valid_a, int nchan_a, int chan_a,
                      const uint32_t* d_words_b, const uint8_t* d_valid_b, int nchan_b, int chan_b, size_t ntime,
                      int n, int overlap, int integrate_ffts, int window, int continuous, fx_cf32* d_out,
                      long* nspec_out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(d_words_a && d_words_b && d_out, FX_ERR_ARG, "fx_spectrometer_xpol: null argument");
    FX_REQUIRE(valid_nchan(nchan_a) && valid_nchan(nchan_b) && chan_a >= 0 && chan_a < nchan_a && chan_b >= 0 &&
                   chan_b < nchan_b,
               FX_ERR_ARG, "fx_spectrometer_xpol: channels %d/%d, %d/%d", chan_a, nchan_a, chan_b, nchan_b);
    FX_REQUIRE(is_pow2(n) && n >= 256 && n <= FX_MAX_N, FX_ERR_ARG, "fx_spectrometer_xpol: n=%d", n);
    FX_REQUIRE(integrate_ffts >= 1 && integrate_ffts <= 65535, FX_ERR_ARG,
               "fx_spectrometer_xpol: integrate_ffts=%d", integrate_ffts);
    FX_REQUIRE(overlap >= 1 && n % overlap == 0, FX_ERR_ARG, "fx_spectrometer_xpol: overlap %d", overlap);
    FX_REQUIRE(window >= FX_WIN_NONE && window <= FX_WIN_HANN, FX_ERR_ARG,
               "fx_spectrometer_xpol: window %d", window);
    const long hop = n / overlap;
    const long K = integrate_ffts;
    long nspec, spec_stride;
    if (continuous) {
        const long nfft = ntime >= (size_t)n ? (long)((ntime - (size_t)n) / (size_t)hop) + 1 : 0;
        nspec = nfft / K;
        spec_stride = K * hop;
    } else {
        spec_stride = (K - 1) * hop + n;
        nspec = (long)(ntime / (size_t)spec_stride);
    }
    if (nspec_out) *nspec_out = nspec;
    if (nspec == 0) return FX_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    float* win = get_window(ctx, window, n);
    FX_REQUIRE(win != nullptr, FX_ERR_NOMEM, "fx_spectrometer_xpol: window table");
    // whole spectra per pass: real frames and their n/2+1-bin transforms of
    // both streams in the context scratch, at most 1 GiB and 65535 frames
    const int c1 = n / 2 + 1;
    const size_t frame_bytes = 2 * ((size_t)n * sizeof(float) + (size_t)c1 * sizeof(float2));
    const size_t budget = (size_t)1 << 30;
    FX_REQUIRE((size_t)K * frame_bytes <= budget, FX_ERR_UNSUPPORTED,
               "fx_spectrometer_xpol: %ld FFTs of %d points per spectrum exceed one pass", K, n);
    long sb = std::min<long>(nspec, (long)(budget / ((size_t)K * frame_bytes)));
    sb = std::max<long>(1, std::min<long>(sb, 65535 / K));
    const size_t nf = (size_t)(sb * K);
    // partial sums: groups(ns) * ns * c1 <= 2^18 + ns * c1 for every pass of ns <= sb spectra
    const size_t part_bytes = ((size_t)262144 + (size_t)sb * (size_t)c1) * sizeof(float2);
    rc = ensure_scratch(ctx, nf * frame_bytes + part_bytes);
    if (rc) return rc;
    float* fa = static_cast<float*>(ctx->scratch);
    float* fb = fa + nf * (size_t)n;
    float2* xa = reinterpret_cast<float2*>(fb + nf * (size_t)n);
    float2* xb = xa + nf * (size_t)c1;
    float2* part = xb + nf * (size_t)c1;
    for (long s0 = 0; s0 < nspec; s0 += sb) {
        const long ns = std::min<long>(sb, nspec - s0);
        const long frames = ns * K;
        FX_CHECK_LAUNCH(fxk_spec_frames(ctx->stream, d_words_a, d_valid_a, nchan_a, chan_a, hop, spec_stride, (int)K, n,
                                        s0 * K, frames, win, fa));
        FX_CHECK_LAUNCH(fxk_spec_frames(ctx->stream, d_words_b, d_valid_b, nchan_b, chan_b, hop, spec_stride, (int)K, n,
                                        s0 * K, frames, win, fb));
        FX_CHECK_LAUNCH(fxk_fft_r2c(ctx->stream, ctx->tw, fa, xa, n, frames, 1));
        FX_CHECK_LAUNCH(fxk_fft_r2c(ctx->stream, ctx->tw, fb, xb, n, frames, 1));
        const int gr = fxk_xpol_groups(c1, (int)K, ns);
        FX_CHECK_LAUNCH(fxk_xpol_acc(ctx->stream, xa, xb, c1, (int)K, ns, 1.0f / (float)K, part, gr,
                                     reinterpret_cast<float2*>(d_out) + (size_t)s0 * c1));
    }
    return FX_OK;
}

// ---------------------------------------------------------------------------
// phase-cal
// ---------------------------------------------------------------------------
struct fx_pcal {
    fx_ctx* ctx;
    int kind;             // 0 trivial, 1 implicit shift, 2 shifting
    double fs, spacing, offset;
    int nbins, ntones, rotatorlen, nfold, step, offset_bin;
    double dphi;
    int ninp = 1;            // inputs sharing this configuration
    double* acc = nullptr;   // ninp x nfold FP64
    double* tones = nullptr; // ninp x 2*ntones FP64
    size_t position = 0;
    uint64_t samplecount = 0;
    bool finalized = false;
};

static long long gcd_ll(long a, long b)
{
    for (;;) {
        a = a % b;
        if (a == 0) return b;
        b = b % a;
        if (b == 0) return a;
    }
}

extern "C" int fx_pcal_create(fx_ctx* ctx, double bandwidth_hz, double spacing_hz, int offset_hz,
                              size_t sampleoffset, fx_pcal** out)
{
    return fx_pcal_create_inputs(ctx, bandwidth_hz, spacing_hz, offset_hz, sampleoffset, 1, out);
}

extern "C" int fx_pcal_create_inputs(fx_ctx* ctx, double bandwidth_hz, double spacing_hz, int offset_hz,
                                     size_t sampleoffset, int ninp, fx_pcal** out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(out && bandwidth_hz > 0 && spacing_hz > 0 && offset_hz >= 0 && ninp >= 1, FX_ERR_ARG,
               "fx_pcal_create: invalid parameters");
    fx_pcal* p = new fx_pcal;
    p->ctx = ctx;
    p->ninp = ninp;
    double fs = 2.0 * bandwidth_hz;
    p->fs = fs;
    p->spacing = spacing_hz;
    p->offset = offset_hz;
    p->rotatorlen = 0;
    if (offset_hz == 0) {
        p->kind = 0;
        p->nbins = (int)(fs / (double)gcd_ll((long)spacing_hz, (long)fs));
        p->ntones = (int)floor(bandwidth_hz / spacing_hz);
    } else {
        int No = (int)(fs / (double)gcd_ll(offset_hz, (long)fs));
        int Np = (int)(fs / (double)gcd_ll((long)spacing_hz, (long)fs));
        p->ntones = (int)floor((bandwidth_hz - offset_hz) / spacing_hz) + 1;
        if (No % Np == 0) {
            p->kind = 1;
            p->nbins = No;
        } else {
            p->kind = 2;
            p->nbins = Np;
            p->rotatorlen = No;
        }
    }
    p->step = (int)floor((double)p->nbins * (spacing_hz / fs));
    p->offset_bin = (p->kind == 1) ? (int)floor((double)p->nbins * (double)offset_hz / fs) : 0;
    p->nfold = p->kind == 2 ? (int)((long long)p->rotatorlen / gcd_ll(p->rotatorlen, p->nbins) * p->nbins)
                            : p->nbins;
    p->dphi = 2.0 * M_PI * (-(double)offset_hz / fs);
    if (hipMalloc(&p->acc, sizeof(double) * (size_t)p->nfold * ninp) != hipSuccess ||
        hipMalloc(&p->tones, sizeof(double) * 2 * (size_t)std::max(p->ntones, 1) * ninp) != hipSuccess) {
        hipGetLastError();
        if (p->acc) hipFree(p->acc);
        delete p;
        fx_set_error("fx_pcal_create: allocation failed");
        return FX_ERR_NOMEM;
    }
    p->position = sampleoffset % (size_t)p->nfold;
    if (hipMemsetAsync(p->acc, 0, sizeof(double) * (size_t)p->nfold * ninp, ctx->stream) != hipSuccess) {
        hipFree(p->acc);
        hipFree(p->tones);
        delete p;
        fx_set_error("fx_pcal_create: clear failed");
        return FX_ERR_HIP;
    }
    *out = p;
    return FX_OK;
}

extern "C" int fx_pcal_destroy(fx_pcal* p)
{
    if (!p) return FX_OK;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
    hipFree(p->acc);
    hipFree(p->tones);
    delete p;
    return FX_OK;
}

extern "C" int fx_pcal_clear(fx_pcal* p)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    int rc = ctx_device(p->ctx);
    if (rc) return rc;
    // fill kernel rather than a memset node (graph-capturable)
    FX_CHECK_LAUNCH(fxk_fill32(p->ctx->stream, p->acc, 0u, 2 * (size_t)p->nfold * p->ninp));
    p->samplecount = 0;
    p->finalized = false;
    return FX_OK;
}

extern "C" int fx_pcal_final_async(fx_pcal* p)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    int rc = ctx_device(p->ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_pcal_final(p->ctx->stream, p->acc, p->nfold, p->nbins, p->ntones, p->offset_bin,
                                   p->step, p->rotatorlen, p->dphi, p->tones, p->ninp));
    p->finalized = true;
    return FX_OK;
}

int fx_pcal_device_state(fx_pcal* p, fx_psub* S, size_t* position)
{
    FX_REQUIRE(p && S, FX_ERR_ARG, "fx_pcal_device_state: null argument");
    S->pacc = p->acc;
    S->tones = p->tones;
    S->nfold = p->nfold;
    S->nbins = p->nbins;
    S->ntones = p->ntones;
    S->offset_bin = p->offset_bin;
    S->step = p->step;
    S->rotatorlen = p->rotatorlen;
    S->dphi = p->dphi;
    if (position) *position = p->position;
    p->finalized = true;   // the caller's kernels fold and extract into acc and tones
    return FX_OK;
}

extern "C" int fx_pcal_tones_input(fx_pcal* p, int input, fx_cf64* tones, uint64_t* nsamples)
{
    FX_REQUIRE(p && tones && input >= 0 && input < p->ninp, FX_ERR_ARG, "fx_pcal_tones: bad argument");
    int rc = ctx_device(p->ctx);
    if (rc) return rc;
    FX_REQUIRE(p->finalized, FX_ERR_STATE, "fx_pcal_tones: no tones extracted yet");
    if (p->ntones > 0)
        FX_HIP_TRY(hipMemcpyAsync(tones, p->tones + (size_t)input * 2 * (size_t)p->ntones,
                                  sizeof(double) * 2 * (size_t)p->ntones, hipMemcpyDeviceToHost, p->ctx->stream));
    FX_HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (nsamples) *nsamples = p->samplecount;
    return FX_OK;
}

extern "C" int fx_pcal_tones(fx_pcal* p, fx_cf64* tones, uint64_t* nsamples)
{
    return fx_pcal_tones_input(p, 0, tones, nsamples);
}

extern "C" int fx_pcal_extract_2bit_inputs(fx_pcal* p, const uint32_t* d_words, size_t in_stride_words,
                                           size_t nsamp)
{
    FX_REQUIRE(p && d_words, FX_ERR_ARG, "null argument");
    FX_REQUIRE(!p->finalized, FX_ERR_STATE, "fx_pcal_extract_2bit: extractor finalized, clear first");
    int rc = ctx_device(p->ctx);
    if (rc) return rc;
    FX_CHECK_LAUNCH(fxk_pcal_fold(p->ctx->stream, d_words, nsamp, p->position, p->nfold, p->acc, p->ninp,
                                  in_stride_words));
    p->position = (p->position + nsamp) % (size_t)p->nfold;
    p->samplecount += nsamp;
    return FX_OK;
}

extern "C" int fx_pcal_adjust_sample_offset(fx_pcal* p, size_t sampleoffset)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    p->position = sampleoffset % (size_t)p->nfold;
    return FX_OK;
}

extern "C" int fx_pcal_extract_2bit(fx_pcal* p, const uint32_t* d_words, size_t nsamp)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    FX_REQUIRE(p->ninp == 1, FX_ERR_ARG, "fx_pcal_extract_2bit: %d inputs, use fx_pcal_extract_2bit_inputs",
               p->ninp);
    return fx_pcal_extract_2bit_inputs(p, d_words, 0, nsamp);
}

extern "C" int fx_pcal_final(fx_pcal* p, fx_cf64* tones, uint64_t* nsamples)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    int rc = ctx_device(p->ctx);
    if (rc) return rc;
    if (!p->finalized) {
        FX_CHECK_LAUNCH(fxk_pcal_final(p->ctx->stream, p->acc, p->nfold, p->nbins, p->ntones,
                                       p->offset_bin, p->step, p->rotatorlen, p->dphi, p->tones, p->ninp));
        p->finalized = true;
    }
    if (tones && p->ntones > 0) {
        FX_HIP_TRY(hipMemcpyAsync(tones, p->tones, sizeof(double) * 2 * (size_t)p->ntones * p->ninp,
                                  hipMemcpyDeviceToHost, p->ctx->stream));
    }
    FX_HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (nsamples) *nsamples = p->samplecount;
    return FX_OK;
}

extern "C" int fx_pcal_info(fx_pcal* p, int* kind, int* nbins, int* ntones)
{
    FX_REQUIRE(p, FX_ERR_ARG, "null pcal");
    if (kind) *kind = p->kind;
    if (nbins) *nbins = p->nbins;
    if (ntones) *ntones = p->ntones;
    return FX_OK;
}

// ---------------------------------------------------------------------------
// correlator
// ---------------------------------------------------------------------------
struct fx_corr {
    fx_ctx* ctx;
    fx_corr_config cfg;
    int nchan;
    size_t elem;            // bytes per channelized complex sample
    void* chan = nullptr;   // [c][t][i], chunk_spectra deep
    void* nat = nullptr;    // F-stage rows of one sub-chunk before the corner turn
    long sub = 0;           // spectra per F-stage sub-chunk (nat depth)
    size_t nat_bytes = 0;   // bytes of the row buffer
    bool direct = true;     // F kernel writes the engine layout itself (no corner turn)
    int nwg = 0;            // workgroups of the persistent direct F kernel (one per CU)
    unsigned* qctr = nullptr;  // its per-XCD block counters, then the X-engine's lockstep words
    int fmode = FX_FSTAGE_AUTO;  // direct F kernel form (fx_corr_set_fstage)
    float* acc = nullptr;   // engine-layout accumulators
    size_t acc_floats = 0;
    fx_phase_model* pm = nullptr;
    const float2* mult = nullptr;
    float* sc_chan = nullptr;
    float* sc_inp = nullptr;
    bool sc_uniform = true;  // every channel scale equals sc_uni (the F kernel folds it per item)
    float sc_uni = 1.0f;
    int2* pij = nullptr;
    long long* coef = nullptr;  // FP8_2BIT / FP4_2BIT: integer coefficients of the exact combination
    int nsets = 1;
    int xtile = 128;        // X-engine input tile (128, or 64 for FP32/FP16 engines)
    bool ring = true;       // the LDS-ring X-engines may run (FX_XENGINE_AUTO, FX_XENGINE_LOCKSTEP)
    bool xlock = false;     // the large-input ring engine may run in lockstep groups (FX_XENGINE_LOCKSTEP)
    long spectra_done = 0;
};

static bool is_twobit(int prec) { return prec == FX_PREC_FP8_2BIT || prec == FX_PREC_FP4_2BIT; }

// engine dispatch: FP32/FP16 read [c][t][i], the 8-bit engines [c][i][t]
static int corr_engine(fx_corr* c, const void* d_chan, long nt)
{
    if (c->cfg.prec == FX_PREC_F32 || c->cfg.prec == FX_PREC_F16)
        return fxk_xengine(c->ctx->stream, c->cfg.prec, c->xtile, d_chan, c->nchan, nt, c->cfg.ninp, c->acc,
                           c->ring, c->xlock && c->qctr ? c->qctr + 8 : nullptr);
    return fxk_xengine8(c->ctx->stream, c->cfg.prec, d_chan, c->nchan, nt, c->cfg.ninp, c->acc);
}

static size_t prec_elem(int prec)
{
    switch (prec) {
    case FX_PREC_F32: return 8;
    case FX_PREC_F16: return 4;
    case FX_PREC_FP8: return 2;
    case FX_PREC_FP8_2BIT: return 1;
    case FX_PREC_FP4_2BIT: return 1;
    default: return 0;
    }
}

extern "C" int fx_corr_create(fx_ctx* ctx, const fx_corr_config* cfg, fx_corr** out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(cfg && out, FX_ERR_ARG, "fx_corr_create: null argument");
    FX_REQUIRE(cfg->ninp >= 1 && cfg->ninp <= 4096, FX_ERR_ARG, "fx_corr_create: ninp=%d", cfg->ninp);
    FX_REQUIRE(is_pow2(cfg->nfft) && cfg->nfft >= 256 && cfg->nfft <= 16384, FX_ERR_ARG,
               "fx_corr_create: nfft=%d not a power of two in [256,16384]", cfg->nfft);
    FX_REQUIRE(cfg->prec >= FX_PREC_F32 && cfg->prec <= FX_PREC_FP4_2BIT, FX_ERR_UNSUPPORTED,
               "fx_corr_create: precision %d", cfg->prec);
    const bool twobit = is_twobit(cfg->prec);
    FX_REQUIRE(cfg->chunk_spectra >= 8 && cfg->chunk_spectra % (twobit ? 16 : 8) == 0, FX_ERR_ARG,
               "fx_corr_create: chunk_spectra=%ld must be a positive multiple of %d", cfg->chunk_spectra,
               twobit ? 16 : 8);
    fx_corr* c = new fx_corr;
    c->ctx = ctx;
    c->cfg = *cfg;
    c->nchan = cfg->nfft / 2;
    c->elem = prec_elem(cfg->prec);
    c->nsets = twobit ? 3 : 1;
    // 2-bit channelized data arrives ready for the engine: no F-stage buffers
    size_t chan_bytes =
        twobit ? 0 : (size_t)c->nchan * (size_t)cfg->chunk_spectra * (size_t)cfg->ninp * c->elem;
    // The F kernel writes natural rows of a sub-chunk of about 64 MB, which the
    // corner turn reads back while they are still in the 256 MB Infinity Cache.
    const size_t row_bytes = (size_t)c->nchan * (size_t)cfg->ninp * c->elem;   // one spectrum of all inputs
    // (FX_FSTAGE_SUB_MB overrides the 64 MB target, for tuning)
    const char* sub_env = getenv("FX_FSTAGE_SUB_MB");
    const size_t sub_target = (size_t)(sub_env ? std::max(1, atoi(sub_env)) : 64) << 20;
    c->sub = std::max<long>(8, std::min<long>(cfg->chunk_spectra, (long)(sub_target / row_bytes) / 8 * 8));
    const char* direct_env = getenv("FX_FSTAGE_DIRECT");
    c->direct = !direct_env || atoi(direct_env) != 0;
    // the direct F kernel's default form: FX_FSTAGE_PERSIST=0 launches it one
    // block per workgroup, FX_FSTAGE_PAIR=0 keeps its persistent blocks single
    const char* persist_env = getenv("FX_FSTAGE_PERSIST");
    const char* pair_env = getenv("FX_FSTAGE_PAIR");
    c->fmode = (persist_env && atoi(persist_env) == 0) ? FX_FSTAGE_BLOCKS
             : (pair_env && atoi(pair_env) == 0)       ? FX_FSTAGE_PERSISTENT
                                                       : FX_FSTAGE_AUTO;
    if (hipDeviceGetAttribute(&c->nwg, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
        c->nwg = 0;
    c->nat_bytes = twobit ? 0 : row_bytes * (size_t)c->sub;
    const size_t nat_bytes = c->nat_bytes;
    // FP32/FP16 engines: 64 x 64 tiles for at most 64 inputs (a 128 x 128 tile
    // would be at least three quarters padding there: 4x the MFMA work)
    c->xtile = (!twobit && cfg->prec != FX_PREC_FP8 && cfg->ninp <= 64) ? 64 : 128;
    c->ring = fxk_xengine_ring_default();
    c->xlock = fxk_xengine_lock_default();
    c->acc_floats = fxk_xengine_acc_floats(c->nchan, cfg->ninp, c->nsets, c->xtile);
    long P = (long)cfg->ninp * (cfg->ninp + 1) / 2;
    std::vector<int2> hp((size_t)P);
    for (int i = 0, p = 0; i < cfg->ninp; i++)
        for (int j = i; j < cfg->ninp; j++, p++) hp[(size_t)p] = make_int2(i, j);
    bool ok = hipMalloc(&c->acc, c->acc_floats * sizeof(float)) == hipSuccess &&
              hipMalloc(&c->pm, sizeof(fx_phase_model) * (size_t)cfg->ninp) == hipSuccess &&
              hipMalloc(&c->pij, sizeof(int2) * (size_t)P) == hipSuccess;
    if (ok && c->nwg >= 8)
        ok = hipMalloc(&c->qctr, (8 + (size_t)fxk_xengine_prog_words()) * sizeof(unsigned)) == hipSuccess;
    if (ok && chan_bytes)
        ok = hipMalloc(&c->chan, chan_bytes) == hipSuccess && hipMalloc(&c->nat, nat_bytes) == hipSuccess;
    if (ok) ok = hipMemcpy(c->pij, hp.data(), sizeof(int2) * (size_t)P, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = hipMemsetAsync(c->acc, 0, c->acc_floats * sizeof(float), ctx->stream) == hipSuccess;
    if (ok) ok = hipMemsetAsync(c->pm, 0, sizeof(fx_phase_model) * (size_t)cfg->ninp, ctx->stream) == hipSuccess;
    if (ok && twobit) {
        // X X^H = (2-H) S S^H + (H-1) U U^H + (H-1)(H-2) T T^H with H = h 2^-22
        // (h an integer: the FP32 level value has 22 fraction bits): 2^44 X X^H
        // = 2^22 [(2^23 - h) SS + (h - 2^22) UU] + (h - 2^22)(h - 2^23) TT
        const double hd = ldexp((double)FX_LUT_HI, 22);
        const long long h = (long long)hd;
        if ((double)h != hd) {
            fx_set_error("fx_corr_create: the level value is not a multiple of 2^-22");
            ok = false;
        }
        const long long c2 = (h - (1LL << 22)) * (h - (1LL << 23));
        const long long w[4] = {(1LL << 23) - h, h - (1LL << 22), c2 >> 22, c2 & ((1LL << 22) - 1)};
        ok = ok && hipMalloc(&c->coef, sizeof(w)) == hipSuccess &&
             hipMemcpy(c->coef, w, sizeof(w), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (ok && (cfg->prec == FX_PREC_F16 || cfg->prec == FX_PREC_FP8)) {
        // default power scaling: a 2-bit stream through the level table has
        // E[x^2] = 0.6726 + 0.3274*H^2 ~ 4.316, so the FFT output has
        // sigma ~ sqrt(N * 4.316); map that to 1 before an FP16 cast and to
        // 16 before an e4m3 cast (normal range 2^-6 .. 448: below 1e-3 sigma
        // and above 28 sigma the format loses relative precision)
        const double target = cfg->prec == FX_PREC_FP8 ? 16.0 : 1.0;
        std::vector<float> sc((size_t)c->nchan, (float)(target / sqrt((double)cfg->nfft * 4.316)));
        c->sc_uniform = true;
        c->sc_uni = sc[0];
        std::vector<float> si((size_t)cfg->ninp, 1.0f);
        ok = hipMalloc(&c->sc_chan, sizeof(float) * sc.size()) == hipSuccess &&
             hipMalloc(&c->sc_inp, sizeof(float) * si.size()) == hipSuccess &&
             hipMemcpy(c->sc_chan, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(c->sc_inp, si.data(), sizeof(float) * si.size(), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        hipGetLastError();
        fx_corr_destroy(c);
        fx_set_error("fx_corr_create: device allocation failed (channel buffer %zu bytes)", chan_bytes);
        return FX_ERR_NOMEM;
    }
    *out = c;
    return FX_OK;
}

extern "C" int fx_corr_destroy(fx_corr* c)
{
    if (!c) return FX_OK;
    hipSetDevice(c->ctx->device);
    hipStreamSynchronize(c->ctx->stream);
    if (c->chan) hipFree(c->chan);
    if (c->nat) hipFree(c->nat);
    if (c->acc) hipFree(c->acc);
    if (c->pm) hipFree(c->pm);
    if (c->pij) hipFree(c->pij);
    if (c->sc_chan) hipFree(c->sc_chan);
    if (c->sc_inp) hipFree(c->sc_inp);
    if (c->coef) hipFree(c->coef);
    if (c->qctr) hipFree(c->qctr);
    delete c;
    return FX_OK;
}

extern "C" int fx_corr_set_phase_model(fx_corr* c, const fx_phase_model* pm)
{
    FX_REQUIRE(c && pm, FX_ERR_ARG, "fx_corr_set_phase_model: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_HIP_TRY(hipMemcpyAsync(c->pm, pm, sizeof(fx_phase_model) * (size_t)c->cfg.ninp,
                              hipMemcpyHostToDevice, c->ctx->stream));
    FX_HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    return FX_OK;
}

extern "C" int fx_corr_set_multipliers(fx_corr* c, const fx_cf32* d_mult)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    c->mult = (const float2*)d_mult;
    return FX_OK;
}

extern "C" int fx_corr_set_scales(fx_corr* c, const float* chan_scale, const float* inp_scale)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    if (!c->sc_chan) {
        if (hipMalloc(&c->sc_chan, sizeof(float) * (size_t)c->nchan) != hipSuccess ||
            hipMalloc(&c->sc_inp, sizeof(float) * (size_t)c->cfg.ninp) != hipSuccess) {
            hipGetLastError();
            fx_set_error("fx_corr_set_scales: allocation failed");
            return FX_ERR_NOMEM;
        }
    }
    std::vector<float> sc((size_t)c->nchan, 1.0f), si((size_t)c->cfg.ninp, 1.0f);
    if (chan_scale) memcpy(sc.data(), chan_scale, sizeof(float) * sc.size());
    if (inp_scale) memcpy(si.data(), inp_scale, sizeof(float) * si.size());
    for (float v : sc) FX_REQUIRE(v > 0.f && std::isfinite(v), FX_ERR_ARG, "scales must be positive");
    for (float v : si) FX_REQUIRE(v > 0.f && std::isfinite(v), FX_ERR_ARG, "scales must be positive");
    FX_HIP_TRY(hipMemcpy(c->sc_chan, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice));
    FX_HIP_TRY(hipMemcpy(c->sc_inp, si.data(), sizeof(float) * si.size(), hipMemcpyHostToDevice));
    c->sc_uniform = std::all_of(sc.begin(), sc.end(), [&](float v) { return v == sc[0]; });
    c->sc_uni = c->sc_uniform ? sc[0] : 1.0f;
    return FX_OK;
}

// F kernel of nspectra (<= sub) spectra into the row buffer nat_buf, on stream s
static int fstage_rows_into(fx_corr* c, const uint32_t* d_words, size_t in_stride_words, long t_first,
                            long nspectra, int out_prec, void* nat_buf, hipStream_t s)
{
    const bool scaled = out_prec != FX_PREC_F32;
    FX_CHECK_LAUNCH(fxk_fstage(s, c->ctx->tw, d_words, in_stride_words, c->cfg.ninp, c->cfg.nfft, t_first,
                               nspectra, c->cfg.use_phase ? c->pm : nullptr, c->mult,
                               scaled && !c->sc_uniform ? c->sc_chan : nullptr,
                               scaled && c->sc_uniform ? c->sc_uni : 1.0f, scaled ? c->sc_inp : nullptr, out_prec,
                               nat_buf, nullptr));
    return FX_OK;
}

// corner turn of the rows in nat_buf (spectra s0 .. s0+nspectra of a
// chunk_spectra-deep engine buffer d_chan), on stream s
static int cornerturn_from(fx_corr* c, long nspectra, long s0, long chunk_spectra, int out_prec, void* d_chan,
                           const void* nat_buf, hipStream_t s)
{
    const long I = c->cfg.ninp;
    const bool imajor = out_prec == FX_PREC_FP8;
    FX_CHECK_LAUNCH(fxk_cornerturn(s, out_prec, nat_buf, d_chan, nspectra * I, c->nchan, chunk_spectra * I,
                                   imajor ? s0 : s0 * I, (int)nspectra, chunk_spectra));
    return FX_OK;
}

static bool fstage_persistent(const fx_corr* c, int out_prec);

extern "C" int fx_corr_fstage(fx_corr* c, const uint32_t* d_words, size_t in_stride_words,
                              long t_first, long nspectra, void* d_chan, int out_prec)
{
    FX_REQUIRE(c && d_words && d_chan, FX_ERR_ARG, "fx_corr_fstage: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_REQUIRE(in_stride_words % 2 == 0 && ((uintptr_t)d_words % 8) == 0, FX_ERR_ARG,
               "fx_corr_fstage: input stride must be even and the base 8-byte aligned");
    FX_REQUIRE(t_first >= 0 && nspectra >= 0, FX_ERR_ARG, "fx_corr_fstage: negative range");
    FX_REQUIRE(nspectra <= c->cfg.chunk_spectra, FX_ERR_ARG,
               "fx_corr_fstage: %ld spectra exceed the %ld-spectrum chunk", nspectra, c->cfg.chunk_spectra);
    FX_REQUIRE(c->nat != nullptr && !is_twobit(out_prec), FX_ERR_UNSUPPORTED,
               "fx_corr_fstage: no F-stage for 2-bit channelized engines");
    FX_REQUIRE(prec_elem(out_prec) > 0 && prec_elem(out_prec) <= c->elem, FX_ERR_ARG,
               "fx_corr_fstage: output precision %d wider than the engine's", out_prec);
    FX_REQUIRE(out_prec != FX_PREC_FP8 || nspectra % 8 == 0, FX_ERR_ARG,
               "fx_corr_fstage: FP8 output needs a multiple of 8 spectra");
    if (nspectra == 0) return FX_OK;
    const long I = c->cfg.ninp;
    const bool imajor = out_prec == FX_PREC_FP8;
    if (c->direct && fxk_fstage_direct_ok(c->cfg.nfft, (int)I, nspectra, out_prec, nspectra * I, 0, nspectra)) {
        const bool scaled = out_prec != FX_PREC_F32;
        FX_CHECK_LAUNCH(fxk_fstage_direct(c->ctx->stream, c->ctx->tw, d_words, in_stride_words, (int)I,
                                          c->cfg.nfft, t_first, nspectra, c->cfg.use_phase ? c->pm : nullptr,
                                          c->mult, scaled && !c->sc_uniform ? c->sc_chan : nullptr,
                                          scaled && c->sc_uniform ? c->sc_uni : 1.0f, scaled ? c->sc_inp : nullptr,
                                          out_prec, d_chan, nspectra * I, 0, imajor ? nspectra : 0,
                                          fstage_persistent(c, out_prec) ? c->qctr : nullptr, c->nwg,
                                          c->fmode == FX_FSTAGE_AUTO));
        return FX_OK;
    }
    // sub-chunks: F kernel into the natural-row buffer, corner turn into d_chan
    for (long s0 = 0; s0 < nspectra; s0 += c->sub) {
        const long ns = std::min(c->sub, nspectra - s0);
        rc = fstage_rows_into(c, d_words, in_stride_words, t_first + s0, ns, out_prec, c->nat, c->ctx->stream);
        if (!rc) rc = cornerturn_from(c, ns, s0, nspectra, out_prec, d_chan, c->nat, c->ctx->stream);
        if (rc) return rc;
    }
    return FX_OK;
}

extern "C" long fx_corr_fstage_depth(fx_corr* c) { return c ? c->sub : 0; }

// the direct F kernel runs in its persistent form (a counter clear, then the
// kernel) for FP32/FP16 output and at most FX_PDIR_PM = 256 inputs
static bool fstage_persistent(const fx_corr* c, int out_prec)
{
    return c->fmode != FX_FSTAGE_BLOCKS && c->qctr && c->nwg >= 8 && c->cfg.ninp <= 256 && out_prec != FX_PREC_FP8;
}

extern "C" int fx_corr_set_fstage(fx_corr* c, int kind)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    FX_REQUIRE(kind == FX_FSTAGE_AUTO || kind == FX_FSTAGE_PERSISTENT || kind == FX_FSTAGE_BLOCKS, FX_ERR_ARG,
               "fx_corr_set_fstage: kind %d", kind);
    c->fmode = kind;
    return FX_OK;
}

extern "C" int fx_corr_fstage_launches(fx_corr* c, long nspectra, int out_prec)
{
    if (!c || nspectra <= 0) return 0;
    const long I = c->cfg.ninp;
    if (c->direct && fxk_fstage_direct_ok(c->cfg.nfft, (int)I, nspectra, out_prec, nspectra * I, 0, nspectra))
        return fstage_persistent(c, out_prec) ? 2 : 1;
    return (int)(2 * ((nspectra + c->sub - 1) / c->sub));
}

extern "C" int fx_corr_fstage_name(fx_corr* c, long nspectra, int out_prec, char* buf, int len)
{
    FX_REQUIRE(c && buf && len > 0, FX_ERR_ARG, "fx_corr_fstage_name: null argument");
    const long I = c->cfg.ninp;
    const int M = c->cfg.nfft / 2;
    const char* t = out_prec == FX_PREC_F32 ? "float2" : out_prec == FX_PREC_F16 ? "__half2" : "fp8x2";
    if (c->direct && nspectra > 0 &&
        fxk_fstage_direct_ok(c->cfg.nfft, (int)I, nspectra, out_prec, nspectra * I, 0, nspectra))
        snprintf(buf, (size_t)len, "%s<%d, %s>%s",
                 fstage_persistent(c, out_prec) ? "k_fstage_direct_p" : "k_fstage_direct", M, t,
                 fstage_persistent(c, out_prec) && c->fmode == FX_FSTAGE_AUTO && out_prec == FX_PREC_F16 &&
                         I % (2 * (16384 / M)) == 0
                     ? " (paired)"
                     : "");
    else
        snprintf(buf, (size_t)len, "k_fstage<%d, %s> + k_cornerturn", M, t);
    return FX_OK;
}

extern "C" int fx_corr_fstage_rows(fx_corr* c, const uint32_t* d_words, size_t in_stride_words,
                                   long t_first, long nspectra, int out_prec)
{
    FX_REQUIRE(c && d_words, FX_ERR_ARG, "fx_corr_fstage_rows: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_REQUIRE(c->nat != nullptr && prec_elem(out_prec) > 0 && prec_elem(out_prec) <= c->elem &&
                   !is_twobit(out_prec),
               FX_ERR_ARG, "fx_corr_fstage_rows: output precision %d", out_prec);
    FX_REQUIRE(in_stride_words % 2 == 0 && ((uintptr_t)d_words % 8) == 0 && t_first >= 0 && nspectra >= 0 &&
                   nspectra <= c->sub,
               FX_ERR_ARG, "fx_corr_fstage_rows: bad range or alignment (at most %ld spectra)", c->sub);
    if (nspectra == 0) return FX_OK;
    return fstage_rows_into(c, d_words, in_stride_words, t_first, nspectra, out_prec, c->nat, c->ctx->stream);
}

extern "C" int fx_corr_cornerturn(fx_corr* c, long nspectra, long s0, long chunk_spectra, int out_prec,
                                  void* d_chan)
{
    FX_REQUIRE(c && d_chan, FX_ERR_ARG, "fx_corr_cornerturn: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_REQUIRE(c->nat != nullptr && nspectra >= 0 && nspectra <= c->sub && s0 >= 0 &&
                   s0 + nspectra <= chunk_spectra && chunk_spectra <= c->cfg.chunk_spectra,
               FX_ERR_ARG, "fx_corr_cornerturn: bad range");
    if (nspectra == 0) return FX_OK;
    return cornerturn_from(c, nspectra, s0, chunk_spectra, out_prec, d_chan, c->nat, c->ctx->stream);
}

extern "C" int fx_corr_process(fx_corr* c, const uint32_t* d_words, size_t in_stride_words,
                               long t_first, long nspectra)
{
    FX_REQUIRE(c && d_words, FX_ERR_ARG, "fx_corr_process: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_REQUIRE(!is_twobit(c->cfg.prec), FX_ERR_UNSUPPORTED,
               "fx_corr_process: a 2-bit channelized engine takes fx_corr_xengine input");
    FX_REQUIRE(c->cfg.prec != FX_PREC_FP8 || nspectra % 8 == 0, FX_ERR_ARG,
               "fx_corr_process: the FP8 engine needs a multiple of 8 spectra");
    for (long done = 0; done < nspectra;) {
        long nt = std::min(c->cfg.chunk_spectra, nspectra - done);
        rc = fx_corr_fstage(c, d_words, in_stride_words, t_first + done, nt, c->chan, c->cfg.prec);
        if (rc) return rc;
        FX_CHECK_LAUNCH(corr_engine(c, c->chan, nt));
        done += nt;
    }
    c->spectra_done += nspectra;
    return FX_OK;
}

extern "C" int fx_corr_xengine(fx_corr* c, const void* d_chan, long nspectra)
{
    FX_REQUIRE(c && d_chan, FX_ERR_ARG, "fx_corr_xengine: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    if (nspectra <= 0) return nspectra == 0 ? FX_OK : FX_ERR_ARG;
    const int mult = is_twobit(c->cfg.prec) ? 16 : c->cfg.prec == FX_PREC_FP8 ? 8 : 1;
    FX_REQUIRE(nspectra % mult == 0, FX_ERR_ARG, "fx_corr_xengine: %ld spectra, need a multiple of %d",
               nspectra, mult);
    FX_REQUIRE(!is_twobit(c->cfg.prec) || c->spectra_done + nspectra <= (1L << 21), FX_ERR_ARG,
               "fx_corr_xengine: more than 2^21 spectra per dump leave the exact FP32 range");
    FX_CHECK_LAUNCH(corr_engine(c, d_chan, nspectra));
    c->spectra_done += nspectra;
    return FX_OK;
}

extern "C" int fx_corr_dump(fx_corr* c, fx_cf32* d_vis)
{
    FX_REQUIRE(c && d_vis, FX_ERR_ARG, "fx_corr_dump: null argument");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    const bool scaled = c->cfg.prec == FX_PREC_F16 || c->cfg.prec == FX_PREC_FP8;
    FX_CHECK_LAUNCH(fxk_xengine_dump_pij(c->ctx->stream, c->acc, c->nchan, c->cfg.ninp, c->xtile, c->nsets, c->coef,
                                         scaled ? c->sc_chan : nullptr, scaled ? c->sc_inp : nullptr,
                                         c->pij, 0, (float2*)d_vis));
    return FX_OK;
}

extern "C" int fx_corr_reset(fx_corr* c)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    // a fill kernel, not a memset node: the reset may be captured in a graph
    FX_CHECK_LAUNCH(fxk_fill32(c->ctx->stream, c->acc, 0u, c->acc_floats));
    c->spectra_done = 0;
    return FX_OK;
}

extern "C" int fx_corr_tune(fx_corr* c, long sub_spectra, int xtile)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    int rc = ctx_device(c->ctx);
    if (rc) return rc;
    FX_REQUIRE(xtile == 0 || xtile == 128 || (xtile == 64 && c->nsets == 1 && c->cfg.prec <= FX_PREC_F16),
               FX_ERR_ARG, "fx_corr_tune: X tile %d (128, or 64 for the FP32/FP16 engines)", xtile);
    FX_REQUIRE(sub_spectra == 0 || (sub_spectra >= 8 && sub_spectra % 8 == 0), FX_ERR_ARG,
               "fx_corr_tune: sub-chunk of %ld spectra", sub_spectra);
    FX_HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    if (xtile && xtile != c->xtile) {
        const size_t n = fxk_xengine_acc_floats(c->nchan, c->cfg.ninp, c->nsets, xtile);
        float* a = nullptr;
        if (hipMalloc(&a, n * sizeof(float)) != hipSuccess) {
            hipGetLastError();
            fx_set_error("fx_corr_tune: accumulator allocation failed");
            return FX_ERR_NOMEM;
        }
        hipFree(c->acc);
        c->acc = a;
        c->acc_floats = n;
        c->xtile = xtile;
        FX_CHECK_LAUNCH(fxk_fill32(c->ctx->stream, c->acc, 0u, c->acc_floats));
    }
    if (sub_spectra && c->nat) {
        const long sub = std::min(sub_spectra, c->cfg.chunk_spectra);
        const size_t bytes = (size_t)c->nchan * (size_t)c->cfg.ninp * c->elem * (size_t)sub;
        if (bytes > c->nat_bytes) {
            void* nat = nullptr;
            if (hipMalloc(&nat, bytes) != hipSuccess) {
                hipGetLastError();
                fx_set_error("fx_corr_tune: row buffer allocation failed");
                return FX_ERR_NOMEM;
            }
            hipFree(c->nat);
            c->nat = nat;
            c->nat_bytes = bytes;
        }
        c->sub = sub;
    }
    return FX_OK;
}

extern "C" int fx_corr_set_xengine(fx_corr* c, int kind)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    FX_REQUIRE(kind == FX_XENGINE_AUTO || kind == FX_XENGINE_PREFETCH || kind == FX_XENGINE_LOCKSTEP, FX_ERR_ARG,
               "fx_corr_set_xengine: kind %d", kind);
    c->ring = kind != FX_XENGINE_PREFETCH;
    c->xlock = kind == FX_XENGINE_LOCKSTEP;
    return FX_OK;
}

extern "C" int fx_corr_xengine_name(fx_corr* c, const void* d_chan, long nspectra, char* buf, int len)
{
    FX_REQUIRE(c && buf && len > 0, FX_ERR_ARG, "fx_corr_xengine_name: null argument");
    const char* name = c->cfg.prec == FX_PREC_FP8        ? "k_xengine8<1>"
                     : c->cfg.prec == FX_PREC_FP8_2BIT ? "k_xengine8<3>"
                     : c->cfg.prec == FX_PREC_FP4_2BIT ? "k_xengine4"
                                                         : fxk_xengine_name(c->cfg.prec, c->xtile,
                                                                            d_chan ? d_chan : c->chan, c->nchan,
                                                                            nspectra, c->cfg.ninp, c->ring,
                                                                            c->xlock && c->qctr);
    snprintf(buf, (size_t)len, "%s", name);
    return FX_OK;
}

extern "C" int fx_corr_info(fx_corr* c, long* chunk, size_t* chan_bytes, size_t* acc_bytes)
{
    FX_REQUIRE(c, FX_ERR_ARG, "null correlator");
    if (chunk) *chunk = c->cfg.chunk_spectra;
    if (chan_bytes) *chan_bytes = (size_t)c->nchan * (size_t)c->cfg.chunk_spectra * (size_t)c->cfg.ninp * c->elem;
    if (acc_bytes) *acc_bytes = c->acc_floats * sizeof(float);
    return FX_OK;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
extern "C" int fx_mfma_probe(fx_ctx* ctx, int which, float* host_out)
{
    int rc = ctx_device(ctx);
    if (rc) return rc;
    FX_REQUIRE(which >= 0 && which <= 2 && host_out, FX_ERR_ARG, "fx_mfma_probe: bad argument");
    float* d = nullptr;
    FX_HIP_TRY(hipMalloc(&d, 32 * 32 * sizeof(float)));
    FX_HIP_TRY(hipMemsetAsync(d, 0, 32 * 32 * sizeof(float), ctx->stream));
    rc = fxk_mfma_probe(ctx->stream, which, d);
    if (rc == FX_OK) {
        hipMemcpyAsync(host_out, d, 32 * 32 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        hipStreamSynchronize(ctx->stream);
    }
    hipFree(d);
    return rc;
}

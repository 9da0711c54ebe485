// spectrometer.cpp -- the spectrometer's control-block and TaskCore seams
// (include/spectrometer_fx.h) on the device spectrometer kernel.
//
// spectrometer_cb_run() is one wake-up of the SPU program
// (spectrometer_spu.c:97-224): stage the raw bytes in HBM, run the windowed
// overlapped FFT power accumulation, write each spectrum in the SPU's
// (power, 0) pair layout and report the mailbox words.  fx_taskcore_* is the
// TaskCore life cycle (TaskCore.h:36-89) with run() asynchronous on a host
// thread that owns the core's HIP stream, like TaskCoreIPP's worker
// (TaskCoreIPP.cpp:214-352).
#include "fx_internal.h"

#include "celldsp_tasklib.h"
#include "spectrometer_fx.h"

#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace {

bool pow2(unsigned n) { return n && !(n & (n - 1)); }

struct Dev {
    fx_ctx* ctx = nullptr;
    void* words = nullptr;
    void* out = nullptr;
    ~Dev()
    {
        if (ctx) {
            if (words) fx_free(ctx, words);
            if (out) fx_free(ctx, out);
            fx_ctx_destroy(ctx);
        }
    }
};

} // namespace

extern "C" int spectrometer_cb_run(const control_block_t* cb, spectrometer_mbox_t* mbox)
{
    FX_REQUIRE(cb && mbox, -1, "spectrometer_cb_run: null argument");
    memset(mbox, 0, sizeof(*mbox));
    if (cb->terminate_flag) return 0;
    const unsigned N = cb->fft_points;
    FX_REQUIRE(pow2(N) && N >= 256 && N <= 32768, -1, "spectrometer_cb_run: fft_points %u", N);
    FX_REQUIRE(cb->samples_per_rawbyte == 4, -1,
               "spectrometer_cb_run: samples_per_rawbyte %u (2-bit = 4 supported)", cb->samples_per_rawbyte);
    const unsigned ov = cb->overlap_factor ? cb->overlap_factor : 1;
    FX_REQUIRE(N % ov == 0, -1, "spectrometer_cb_run: overlap %u", ov);
    const unsigned K = cb->integrate_ffts;
    FX_REQUIRE(K >= 1, -1, "spectrometer_cb_run: integrate_ffts 0");
    FX_REQUIRE(cb->rawif_src.p && cb->spectrum_dst.p, -1, "spectrometer_cb_run: buffers missing");
    const size_t nwords = cb->rawif_bytecount / 4;
    const size_t ntime = nwords * 16;
    const int C = (int)N / 2;
    Dev d;
    if (fx_ctx_create(0, &d.ctx) || fx_malloc(d.ctx, nwords * 4 + 16, &d.words)) return -1;
    if (fx_memcpy_h2d(d.ctx, d.words, cb->rawif_src.p, nwords * 4)) return -1;
    const size_t hop = N / ov;
    const size_t nfft = ntime >= N ? (ntime - N) / hop + 1 : 0;
    const size_t nspec = nfft / K;
    if (nspec == 0) {
        mbox->total_ffts = (unsigned)nfft;
        return 0;
    }
    if (fx_malloc(d.ctx, nspec * (size_t)(C + 1) * sizeof(float), &d.out)) return -1;
    void *e0 = nullptr, *e1 = nullptr;
    if (fx_event_create(d.ctx, &e0) || fx_event_create(d.ctx, &e1)) return -1;
    long got = 0;
    fx_event_record(d.ctx, e0);
    int rc = fx_spectrometer_ex(d.ctx, (const uint32_t*)d.words, ntime, 1, (int)N, (int)ov, (int)K,
                                FX_WIN_SIN2, 1, 1, (float*)d.out, &got);
    fx_event_record(d.ctx, e1);
    float ms = 0.f;
    if (rc == FX_OK) rc = fx_event_elapsed_ms(e0, e1, &ms);
    fx_event_destroy(e0);
    fx_event_destroy(e1);
    if (rc) return -1;
    std::vector<float> h(nspec * (size_t)(C + 1));
    if (fx_memcpy_d2h(d.ctx, h.data(), d.out, h.size() * sizeof(float))) return -1;
    float* dst = (float*)cb->spectrum_dst.p;
    for (size_t s = 0; s < nspec; s++) {
        const float* p = &h[s * (size_t)(C + 1)];
        float* o = dst + s * 2 * (size_t)N;
        for (unsigned k = 0; k < N; k++) {
            unsigned kk = k <= (unsigned)C ? k : N - k;
            o[2 * k] = p[kk];
            o[2 * k + 1] = 0.f;
        }
    }
    mbox->t_spu = (unsigned)(ms * 1e-3 * CELLDSP_TIMEBASE);
    mbox->num_spectra = (unsigned)nspec;
    mbox->total_ffts = (unsigned)nfft;
    return 0;
}

// ---------------------------------------------------------------------------
// phase-cal of the IPP core, on device-resident words
// ---------------------------------------------------------------------------
// Phase-cal of every spectrum: the FFT inputs whose index (counted from
// spectrum spec0 * K) is a multiple of the overlap factor (TaskCoreIPP.cpp:
// 413-418) are contiguous within a spectrum, so each spectrum is one span of
// whole FFTs.  The span is folded exactly on the device into rotatorlen FP64
// bins (every 2-bit value is a multiple of 2^-22, so the sums are exact and
// order-free); the rotator and the fold into tonebins (TaskCoreIPP.cpp:
// 633-693) are then applied in FP64 on rotatorlen values per spectrum.
int fx_spec_pcal(fx_ctx* ctx, const uint32_t* d_words, long nspec, long L, long spec0, int N, int ov, int K, int R,
                 int T, double dphi, void*& fold, size_t& fold_cap, std::vector<fx_cf64>& out,
                 const uint8_t* d_wvalid)
{
    const long hop = N / ov;
    const size_t bytes = (size_t)nspec * (size_t)R * sizeof(double);
    if (fold_cap < bytes) {
        if (fold) fx_free(ctx, fold);
        fold = nullptr;
        fold_cap = 0;
        if (fx_malloc(ctx, bytes, &fold) != FX_OK) return -1;
        fold_cap = bytes;
    }
    if (fx_memset(ctx, fold, 0, bytes)) return -1;
    hipStream_t st = fx_ctx_hip_stream(ctx);
    for (long sp = 0; sp < nspec; sp++) {
        const long g0 = (spec0 + sp) * (long)K;      // index of the spectrum's first FFT
        const long j0 = (ov - g0 % ov) % ov;         // first selected FFT of this spectrum
        if (j0 >= K) continue;
        const long cnt = (K - 1 - j0) / ov + 1;
        const size_t start = (size_t)(sp * L + j0 * hop);
        if (fxk_pcal_fold(st, d_words + start / 16, (size_t)(cnt * N), 0, R, (double*)fold + (size_t)sp * R, 1, 0,
                          d_wvalid ? d_wvalid + start / 16 : nullptr) != FX_OK) {
            fx_set_error("phase-cal fold launch failed");
            return -1;
        }
    }
    std::vector<double> h((size_t)nspec * R);
    if (fx_memcpy_d2h(ctx, h.data(), fold, bytes)) return -1;
    std::vector<double> rc((size_t)R), rs((size_t)R);
    for (int r = 0; r < R; r++) {
        rc[(size_t)r] = cos(dphi * (double)r);
        rs[(size_t)r] = sin(dphi * (double)r);
    }
    const int pulses = R / T;
    out.assign((size_t)nspec * T, fx_cf64{0.0, 0.0});
    for (long sp = 0; sp < nspec; sp++) {
        const double* f = &h[(size_t)sp * R];
        fx_cf64* o = &out[(size_t)sp * T];
        for (int p = 0; p < pulses; p++)
            for (int b = 0; b < T; b++) {
                const int r = p * T + b;
                o[b].re += f[r] * rc[(size_t)r];
                o[b].im += f[r] * rs[(size_t)r];
            }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// TaskCore
// ---------------------------------------------------------------------------
struct fx_taskcore {
    int rank = 0;
    fx_spectrometer_settings s{};
    fx_ctx* ctx = nullptr;
    // phase-cal of the IPP core (swspectrometer.cpp:178-196)
    int tonebins = 0, rotatorlen = 0;
    double dphi = 0.0;
    void* words = nullptr;
    size_t words_cap = 0;
    void* out = nullptr;
    size_t out_cap = 0;
    void* fold = nullptr;
    size_t fold_cap = 0;
    std::thread th;
    int result = 0;
    long nspec = 0;
    std::vector<float> last;
    std::vector<fx_cf64> last_pcal;
    std::string err;
};

namespace {

int ensure(fx_ctx* ctx, void*& p, size_t& cap, size_t bytes)
{
    if (cap >= bytes) return FX_OK;
    if (p) fx_free(ctx, p);
    p = nullptr;
    cap = 0;
    int rc = fx_malloc(ctx, bytes, &p);
    if (rc == FX_OK) cap = bytes;
    return rc;
}

int core_pcal(fx_taskcore* c, long nspec, long L, fx_cf64* pcal_out)
{
    const fx_spectrometer_settings& s = c->s;
    if (fx_spec_pcal(c->ctx, (const uint32_t*)c->words, nspec, L, 0, s.fft_points, s.fft_overlap_factor,
                     s.core_overlapped_ffts, c->rotatorlen, c->tonebins, c->dphi, c->fold, c->fold_cap,
                     c->last_pcal)) {
        c->err = "phase-cal fold failed";
        return -1;
    }
    memcpy(pcal_out, c->last_pcal.data(), c->last_pcal.size() * sizeof(fx_cf64));
    return 0;
}

int core_work(fx_taskcore* c, const void* raw, size_t rawbytes, float* spectra_out, size_t cap,
              fx_cf64* pcal_out, size_t pcal_cap)
{
    const fx_spectrometer_settings& s = c->s;
    const int N = s.fft_points, ov = s.fft_overlap_factor, K = s.core_overlapped_ffts;
    const int nch = s.source_channels;
    const int C1 = N / 2 + 1;
    const size_t nwords = rawbytes / 4;
    const size_t ntime = nwords * 16 / (size_t)nch;
    c->last.clear();
    c->last_pcal.clear();
    c->nspec = 0;
    // whole spectra only: the IPP core's trailing partial spectrum
    // (TaskCoreIPP.cpp:521-526) is never produced
    const long L = (long)(K - 1) * (N / ov) + N;
    const long nspec = (long)(ntime / (size_t)L);
    if (nspec == 0) return 0;
    const int outch = s.use_channel ? 1 : nch;
    if ((size_t)nspec * outch * C1 > cap) {
        c->err = "spectra_out too small";
        return -1;
    }
    const bool want_pcal = c->tonebins > 0 && pcal_out;
    if (want_pcal && (size_t)nspec * c->tonebins > pcal_cap) {
        c->err = "pcal_out too small";
        return -1;
    }
    if (ensure(c->ctx, c->words, c->words_cap, nwords * 4 + 16)) return -1;
    if (fx_memcpy_h2d(c->ctx, c->words, raw, nwords * 4)) return -1;
    if (ensure(c->ctx, c->out, c->out_cap, (size_t)nspec * nch * C1 * sizeof(float))) return -1;
    long got = 0;
    if (fx_spectrometer_ex(c->ctx, (const uint32_t*)c->words, ntime, nch, N, ov, K, s.window, 0, 1,
                           (float*)c->out, &got))
        return -1;
    std::vector<float> h((size_t)nspec * nch * C1);
    if (fx_memcpy_d2h(c->ctx, h.data(), c->out, h.size() * sizeof(float))) return -1;
    c->last.assign((size_t)nspec * outch * C1, 0.f);
    for (long sp = 0; sp < nspec; sp++)
        for (int o = 0; o < outch; o++) {
            const int ch = s.use_channel ? s.use_channel - 1 : o;
            memcpy(&c->last[((size_t)sp * outch + o) * C1], &h[((size_t)sp * nch + ch) * C1],
                   C1 * sizeof(float));
        }
    memcpy(spectra_out, c->last.data(), c->last.size() * sizeof(float));
    if (want_pcal && core_pcal(c, nspec, L, pcal_out)) return -1;
    c->nspec = nspec;
    return (int)nspec;
}

} // namespace

extern "C" int fx_taskcore_create(int rank, const fx_spectrometer_settings* s, fx_taskcore** out)
{
    FX_REQUIRE(s && out, FX_ERR_ARG, "fx_taskcore_create: null argument");
    *out = nullptr;
    FX_REQUIRE(pow2((unsigned)s->fft_points) && s->fft_points >= 256 && s->fft_points <= 32768, FX_ERR_ARG,
               "fx_taskcore_create: fft_points %d", s->fft_points);
    FX_REQUIRE(s->fft_overlap_factor >= 1 && s->fft_points % s->fft_overlap_factor == 0 &&
                   (s->fft_points / s->fft_overlap_factor) % 16 == 0,
               FX_ERR_ARG, "fx_taskcore_create: overlap %d", s->fft_overlap_factor);
    FX_REQUIRE(s->core_overlapped_ffts >= 1, FX_ERR_ARG, "fx_taskcore_create: core_overlapped_ffts");
    FX_REQUIRE(s->window >= FX_WIN_NONE && s->window <= FX_WIN_HANN, FX_ERR_ARG,
               "fx_taskcore_create: window %d", s->window);
    FX_REQUIRE(s->bits_per_sample == 2, FX_ERR_UNSUPPORTED, "fx_taskcore_create: %d-bit input",
               s->bits_per_sample);
    const int nc = s->source_channels;
    FX_REQUIRE(nc == 1 || nc == 2 || nc == 4 || nc == 8 || nc == 16, FX_ERR_ARG,
               "fx_taskcore_create: source_channels %d", nc);
    FX_REQUIRE(s->use_channel >= 0 && s->use_channel <= nc, FX_ERR_ARG, "fx_taskcore_create: use_channel");
    int T = 0, R = 0;
    if (s->extract_pcal) {
        FX_REQUIRE(nc == 1, FX_ERR_UNSUPPORTED, "fx_taskcore_create: phase-cal on multi-channel sources");
        FX_REQUIRE(s->samplingfreq > 0 && s->pcalharmonicshz > 0 && s->pcaloffsethz >= 0, FX_ERR_ARG,
                   "fx_taskcore_create: phase-cal frequencies");
        const long long fs = (long long)s->samplingfreq, off = (long long)s->pcaloffsethz;
        T = (int)(s->samplingfreq / s->pcalharmonicshz);
        // rotator period fs/gcd(fs, offset); a zero offset needs no rotation,
        // one tonebins period stands in for the reference's gcd(fs, 0)
        R = off ? (int)(fs / std::gcd(fs, off)) : T;
        FX_REQUIRE(T >= 1 && R % T == 0, FX_ERR_UNSUPPORTED,
                   "fx_taskcore_create: phase-cal rotator length %d is not a multiple of %d tone bins", R, T);
        if (s->fft_points % R != 0) {
            // swspectrometer.cpp:185-190: the rotator must fit evenly into an
            // FFT input, else phase-cal extraction is switched off
            T = 0;
            R = 0;
        }
    }
    fx_taskcore* c = new fx_taskcore;
    c->rank = rank;
    c->s = *s;
    c->tonebins = T;
    c->rotatorlen = R;
    c->dphi = 2.0 * M_PI * (-s->pcaloffsethz / s->samplingfreq);
    int rc = fx_ctx_create(0, &c->ctx);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return FX_OK;
}

extern "C" int fx_taskcore_run(fx_taskcore* c, const void* raw, size_t rawbytes, float* spectra_out,
                               size_t cap, fx_cf64* pcal_out, size_t pcal_cap)
{
    FX_REQUIRE(c && raw && spectra_out, FX_ERR_ARG, "fx_taskcore_run: null argument");
    FX_REQUIRE(!c->th.joinable(), FX_ERR_STATE, "fx_taskcore_run: previous run not joined");
    c->err.clear();
    c->result = 0;
    c->th = std::thread([=]() {
        c->result = core_work(c, raw, rawbytes, spectra_out, cap, pcal_out, pcal_cap);
        if (c->result < 0 && c->err.empty()) c->err = fx_last_error();
    });
    return FX_OK;
}

extern "C" int fx_taskcore_join(fx_taskcore* c)
{
    FX_REQUIRE(c, -1, "fx_taskcore_join: null core");
    if (c->th.joinable()) c->th.join();
    if (c->result < 0) {
        fx_set_error("TaskCore %d: %s", c->rank, c->err.empty() ? "device error" : c->err.c_str());
        return -1;
    }
    return c->result;
}

extern "C" int fx_taskcore_combine(fx_taskcore* c, float* out, size_t n, fx_cf64* pcal, size_t npcal)
{
    FX_REQUIRE(c && out, FX_ERR_ARG, "fx_taskcore_combine: null argument");
    FX_REQUIRE(!c->th.joinable(), FX_ERR_STATE, "fx_taskcore_combine: run not joined");
    const size_t m = std::min(n, c->last.size());
    for (size_t k = 0; k < m; k++) out[k] += c->last[k];
    if (pcal) {
        const size_t q = std::min(npcal, c->last_pcal.size());
        for (size_t k = 0; k < q; k++) {
            pcal[k].re += c->last_pcal[k].re;
            pcal[k].im += c->last_pcal[k].im;
        }
    }
    return FX_OK;
}

extern "C" int fx_taskcore_pcal_length(fx_taskcore* c) { return c ? c->tonebins : 0; }

extern "C" int fx_taskcore_finalize(fx_taskcore* c)
{
    FX_REQUIRE(c, FX_ERR_ARG, "fx_taskcore_finalize: null core");
    if (c->th.joinable()) c->th.join();
    return fx_ctx_sync(c->ctx);
}

extern "C" int fx_taskcore_destroy(fx_taskcore* c)
{
    if (!c) return FX_OK;
    if (c->th.joinable()) c->th.join();
    if (c->ctx) {
        if (c->words) fx_free(c->ctx, c->words);
        if (c->out) fx_free(c->ctx, c->out);
        if (c->fold) fx_free(c->ctx, c->fold);
        fx_ctx_destroy(c->ctx);
    }
    delete c;
    return FX_OK;
}

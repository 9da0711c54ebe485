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

#include <climits>
#include <cmath>
#include <cstdint>
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

extern "C" int spectrometer_cb_run(const control_block_t* cb_in, spectrometer_mbox_t* mbox)
{
    FX_REQUIRE(cb_in && mbox, -1, "spectrometer_cb_run: null argument");
    // the type is 128-byte aligned (the reference's SPU DMA needs it), but a
    // host caller -- an FFI binding -- need not align the block: read it as
    // bytes into an aligned copy
    control_block_t cbv;
    memcpy(&cbv, reinterpret_cast<const void*>(reinterpret_cast<uintptr_t>(cb_in)), sizeof(cbv));
    const control_block_t* cb = &cbv;
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
                 const uint8_t* d_wvalid, int kuse)
{
    if (kuse < 0) kuse = K;
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
        if (j0 >= kuse) continue;
        const long cnt = (kuse - 1 - j0) / ov + 1;
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
    // fx_taskcore_run_buffers: device copies of the sources, the last run's
    // spectra per source / pair and phase-cal per source
    int nsrc = 0;
    std::vector<void*> srcw;
    std::vector<size_t> srcw_cap;
    void* xout = nullptr;
    size_t xout_cap = 0;
    std::vector<std::vector<float>> m_auto;
    std::vector<std::vector<fx_cf32>> m_xpol;
    std::vector<std::vector<fx_cf64>> m_pcal;
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

// The IPP core's doMaths over nsources sources (TaskCoreIPP.cpp:356-560);
// returns the spectra written (full + partial) or -1.
int core_work_buffers(fx_taskcore* c, int ns, const std::vector<fx_buffer*>& in, const std::vector<fx_buffer*>& outb,
                      const std::vector<fx_buffer*>& xpb, const std::vector<fx_buffer*>& pcb)
{
    const fx_spectrometer_settings& s = c->s;
    const int N = s.fft_points, ov = s.fft_overlap_factor, K = s.core_overlapped_ffts;
    const int nch = s.source_channels;
    const int C1 = N / 2 + 1;
    const long hop = N / ov;
    const int ch0 = (s.use_channel ? s.use_channel : 1) - 1;
    const int ch1 = s.use_channel_file2 ? s.use_channel_file2 - 1 : ch0;
    auto chan_of = [&](int k) { return k == 0 ? ch0 : ch1; };
    const int np = ns * (ns - 1) / 2;
    c->m_auto.assign((size_t)ns, {});
    c->m_xpol.assign((size_t)(xpb.empty() ? 0 : np), {});
    c->m_pcal.assign((size_t)ns, {});
    // samples per channel every source still has
    size_t ntime = SIZE_MAX;
    for (int k = 0; k < ns; k++) ntime = std::min(ntime, in[(size_t)k]->length / 4 * 16 / (size_t)nch);
    const long L = (long)(K - 1) * hop + N;
    const long nfull = (long)(ntime / (size_t)L);
    const long rest = (long)ntime - nfull * L;
    // whole FFTs left after the last full spectrum (at most K-1)
    const long f = rest >= N ? std::min<long>(K - 1, (rest - N) / hop + 1) : 0;
    const long total = nfull + (f > 0 ? 1 : 0);
    const bool pcal = c->tonebins > 0 && !pcb.empty();
    const int T = c->tonebins;
    for (int k = 0; k < ns; k++) {
        if (outb[(size_t)k]->allocated < (size_t)total * C1 * sizeof(float)) {
            c->err = "outbuf too small for " + std::to_string(total) + " spectra";
            return -1;
        }
        if (pcal && pcb[(size_t)k]->allocated < (size_t)total * T * sizeof(fx_cf32)) {
            c->err = "pcalbuf too small";
            return -1;
        }
    }
    for (int x = 0; x < (int)c->m_xpol.size(); x++)
        if (xpb[(size_t)x]->allocated < (size_t)total * C1 * sizeof(fx_cf32)) {
            c->err = "xpolbuf too small";
            return -1;
        }
    if (total == 0) {
        for (int k = 0; k < ns; k++) {
            outb[(size_t)k]->length = 0;
            if (pcal) pcb[(size_t)k]->length = 0;
        }
        for (size_t x = 0; x < c->m_xpol.size(); x++) xpb[x]->length = 0;
        return 0;
    }
    // stage the samples the spectra use (whole words; every spectrum and FFT
    // start is a multiple of 16 samples, so of nch/16 words)
    const size_t used = (size_t)(nfull * L + (f ? (f - 1) * hop + N : 0));
    const size_t nwords = (used * (size_t)nch + 15) / 16;
    c->srcw.resize((size_t)ns, nullptr);
    c->srcw_cap.resize((size_t)ns, 0);
    for (int k = 0; k < ns; k++) {
        if (ensure(c->ctx, c->srcw[(size_t)k], c->srcw_cap[(size_t)k], nwords * 4 + 16)) return -1;
        if (fx_memcpy_h2d(c->ctx, c->srcw[(size_t)k], in[(size_t)k]->data, nwords * 4)) return -1;
    }
    const size_t tail_off = (size_t)nfull * L * nch / 16;   // words to the partial spectrum
    const size_t tail_t = f ? (size_t)((f - 1) * hop + N) : 0;
    // autos: every channel of each source through the device spectrometer,
    // then its selected channel
    if (ensure(c->ctx, c->out, c->out_cap, (size_t)std::max<long>(nfull, 1) * nch * C1 * sizeof(float))) return -1;
    std::vector<float> h;
    for (int k = 0; k < ns; k++) {
        std::vector<float>& a = c->m_auto[(size_t)k];
        a.assign((size_t)total * C1, 0.f);
        const uint32_t* w = (const uint32_t*)c->srcw[(size_t)k];
        const int ch = chan_of(k);
        if (nfull) {
            long got = 0;
            if (fx_spectrometer_ex(c->ctx, w, (size_t)nfull * L, nch, N, ov, K, s.window, 0, 1, (float*)c->out,
                                   &got))
                return -1;
            h.resize((size_t)nfull * nch * C1);
            if (fx_memcpy_d2h(c->ctx, h.data(), c->out, h.size() * sizeof(float))) return -1;
            for (long sp = 0; sp < nfull; sp++)
                memcpy(&a[(size_t)sp * C1], &h[((size_t)sp * nch + ch) * C1], C1 * sizeof(float));
        }
        if (f) {
            // the partial spectrum holds sums: the mean of f FFTs times f
            long got = 0;
            if (fx_spectrometer_ex(c->ctx, w + tail_off, tail_t, nch, N, ov, (int)f, s.window, 0, 1,
                                   (float*)c->out, &got))
                return -1;
            h.resize((size_t)nch * C1);
            if (fx_memcpy_d2h(c->ctx, h.data(), c->out, h.size() * sizeof(float))) return -1;
            for (int b = 0; b < C1; b++) a[(size_t)nfull * C1 + b] = h[(size_t)ch * C1 + b] * (float)f;
        }
    }
    // cross spectra of every source pair
    if (!c->m_xpol.empty()) {
        if (ensure(c->ctx, c->xout, c->xout_cap, (size_t)std::max<long>(nfull, 1) * C1 * sizeof(fx_cf32))) return -1;
        int x = 0;
        for (int a = 0; a < ns; a++)
            for (int b = a + 1; b < ns; b++, x++) {
                std::vector<fx_cf32>& v = c->m_xpol[(size_t)x];
                v.assign((size_t)total * C1, fx_cf32{0.f, 0.f});
                const uint32_t* wa = (const uint32_t*)c->srcw[(size_t)a];
                const uint32_t* wb = (const uint32_t*)c->srcw[(size_t)b];
                long got = 0;
                if (nfull) {
                    if (fx_spectrometer_xpol(c->ctx, wa, nch, chan_of(a), wb, nch, chan_of(b), (size_t)nfull * L, N, ov,
                                             K, s.window, 0, (fx_cf32*)c->xout, &got))
                        return -1;
                    if (fx_memcpy_d2h(c->ctx, v.data(), c->xout, (size_t)nfull * C1 * sizeof(fx_cf32))) return -1;
                }
                if (f) {
                    if (fx_spectrometer_xpol(c->ctx, wa + tail_off, nch, chan_of(a), wb + tail_off, nch, chan_of(b),
                                             tail_t, N, ov, (int)f, s.window, 0, (fx_cf32*)c->xout, &got))
                        return -1;
                    fx_cf32* t = &v[(size_t)nfull * C1];
                    if (fx_memcpy_d2h(c->ctx, t, c->xout, (size_t)C1 * sizeof(fx_cf32))) return -1;
                    for (int b2 = 0; b2 < C1; b2++) {
                        t[b2].re *= (float)f;
                        t[b2].im *= (float)f;
                    }
                }
            }
    }
    // phase-cal of each source (FP64 accumulation)
    if (pcal) {
        std::vector<fx_cf64> part;
        for (int k = 0; k < ns; k++) {
            const uint32_t* w = (const uint32_t*)c->srcw[(size_t)k];
            std::vector<fx_cf64>& pv = c->m_pcal[(size_t)k];
            pv.assign((size_t)total * T, fx_cf64{0.0, 0.0});
            if (nfull) {
                if (fx_spec_pcal(c->ctx, w, nfull, L, 0, N, ov, K, c->rotatorlen, T, c->dphi, c->fold, c->fold_cap,
                                 part))
                    return -1;
                std::copy(part.begin(), part.end(), pv.begin());
            }
            if (f) {
                if (fx_spec_pcal(c->ctx, w + tail_off, 1, L, nfull, N, ov, K, c->rotatorlen, T, c->dphi, c->fold,
                                 c->fold_cap, part, nullptr, (int)f))
                    return -1;
                std::copy(part.begin(), part.end(), pv.begin() + (size_t)nfull * T);
            }
        }
    }
    // into the caller's buffers
    for (int k = 0; k < ns; k++) {
        memcpy(outb[(size_t)k]->data, c->m_auto[(size_t)k].data(), (size_t)total * C1 * sizeof(float));
        outb[(size_t)k]->length = (size_t)total * C1 * sizeof(float);
        if (pcal) {
            fx_cf32* o = (fx_cf32*)pcb[(size_t)k]->data;
            for (size_t e = 0; e < (size_t)total * T; e++)
                o[e] = fx_cf32{(float)c->m_pcal[(size_t)k][e].re, (float)c->m_pcal[(size_t)k][e].im};
            pcb[(size_t)k]->length = (size_t)total * T * sizeof(fx_cf32);
        }
    }
    for (size_t x = 0; x < c->m_xpol.size(); x++) {
        memcpy(xpb[x]->data, c->m_xpol[x].data(), (size_t)total * C1 * sizeof(fx_cf32));
        xpb[x]->length = (size_t)total * C1 * sizeof(fx_cf32);
    }
    c->nsrc = ns;
    return (int)total;
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
    FX_REQUIRE(s->use_channel >= 0 && s->use_channel <= nc && s->use_channel_file2 >= 0 &&
                   s->use_channel_file2 <= nc,
               FX_ERR_ARG, "fx_taskcore_create: use_channel");
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

extern "C" int fx_taskcore_run_buffers(fx_taskcore* c, int ns, fx_buffer* const* inbuf, fx_buffer* const* outbuf,
                                       fx_buffer* const* xpolbuf, fx_buffer* const* pcalbuf)
{
    FX_REQUIRE(c && inbuf && outbuf, FX_ERR_ARG, "fx_taskcore_run_buffers: null argument");
    FX_REQUIRE(ns >= 1 && ns <= FX_TC_MAX_SOURCES, FX_ERR_ARG, "fx_taskcore_run_buffers: %d sources", ns);
    FX_REQUIRE(!c->th.joinable(), FX_ERR_STATE, "fx_taskcore_run_buffers: previous run not joined");
    const int np = ns * (ns - 1) / 2;
    std::vector<fx_buffer*> in((size_t)ns), out((size_t)ns), xp, pc;
    for (int k = 0; k < ns; k++) {
        FX_REQUIRE(inbuf[k] && outbuf[k] && (inbuf[k]->data || !inbuf[k]->length) && outbuf[k]->data, FX_ERR_ARG,
                   "fx_taskcore_run_buffers: source %d buffers", k);
        in[(size_t)k] = inbuf[k];
        out[(size_t)k] = outbuf[k];
    }
    if (xpolbuf && np) {
        for (int x = 0; x < np; x++) {
            FX_REQUIRE(xpolbuf[x] && xpolbuf[x]->data, FX_ERR_ARG, "fx_taskcore_run_buffers: xpol buffer %d", x);
            xp.push_back(xpolbuf[x]);
        }
    }
    if (pcalbuf && c->tonebins > 0) {
        for (int k = 0; k < ns; k++) {
            FX_REQUIRE(pcalbuf[k] && pcalbuf[k]->data, FX_ERR_ARG, "fx_taskcore_run_buffers: pcal buffer %d", k);
            pc.push_back(pcalbuf[k]);
        }
    }
    c->err.clear();
    c->result = 0;
    c->th = std::thread([=]() {
        c->result = core_work_buffers(c, ns, in, out, xp, pc);
        if (c->result < 0 && c->err.empty()) c->err = fx_last_error();
    });
    return FX_OK;
}

extern "C" int fx_taskcore_combine_buffers(fx_taskcore* c, fx_buffer* const* outbuf, fx_buffer* const* pcalbuf)
{
    FX_REQUIRE(c && outbuf, FX_ERR_ARG, "fx_taskcore_combine_buffers: null argument");
    FX_REQUIRE(!c->th.joinable(), FX_ERR_STATE, "fx_taskcore_combine_buffers: run not joined");
    const int C1 = c->s.fft_points / 2 + 1;
    for (int k = 0; k < c->nsrc; k++) {
        const std::vector<float>& a = c->m_auto[(size_t)k];
        FX_REQUIRE(outbuf[k] && outbuf[k]->allocated >= C1 * sizeof(float), FX_ERR_ARG,
                   "fx_taskcore_combine_buffers: outbuf %d", k);
        float* o = (float*)outbuf[k]->data;
        for (int b = 0; b < C1 && (size_t)b < a.size(); b++) o[b] += a[(size_t)b];
        outbuf[k]->length = C1 * sizeof(float);
    }
    for (size_t x = 0; x < c->m_xpol.size(); x++) {
        fx_buffer* ob = outbuf[(size_t)c->nsrc + x];
        FX_REQUIRE(ob && ob->allocated >= C1 * sizeof(fx_cf32), FX_ERR_ARG,
                   "fx_taskcore_combine_buffers: cross-spectrum outbuf %zu", x);
        fx_cf32* o = (fx_cf32*)ob->data;
        const std::vector<fx_cf32>& v = c->m_xpol[x];
        for (int b = 0; b < C1 && (size_t)b < v.size(); b++) {
            o[b].re += v[(size_t)b].re;
            o[b].im += v[(size_t)b].im;
        }
        ob->length = C1 * sizeof(fx_cf32);
    }
    if (pcalbuf && c->tonebins > 0) {
        const int T = c->tonebins;
        for (int k = 0; k < c->nsrc; k++) {
            const std::vector<fx_cf64>& v = c->m_pcal[(size_t)k];
            if (v.empty()) continue;
            FX_REQUIRE(pcalbuf[k] && pcalbuf[k]->allocated >= T * sizeof(fx_cf32), FX_ERR_ARG,
                       "fx_taskcore_combine_buffers: pcalbuf %d", k);
            fx_cf32* o = (fx_cf32*)pcalbuf[k]->data;
            for (int b = 0; b < T; b++) {
                o[b].re += (float)v[(size_t)b].re;
                o[b].im += (float)v[(size_t)b].im;
            }
            pcalbuf[k]->length = T * sizeof(fx_cf32);
        }
    }
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
        if (c->xout) fx_free(c->ctx, c->xout);
        for (void* w : c->srcw)
            if (w) fx_free(c->ctx, w);
        fx_ctx_destroy(c->ctx);
    }
    delete c;
    return FX_OK;
}

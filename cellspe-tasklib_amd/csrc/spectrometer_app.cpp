// spectrometer_app.cpp -- the software spectrometer's dispatcher loop
// (TaskDispatcher.cpp:38-230, swspectrometer.cpp:150-420) on one GPU.
//
// One or two input files are streamed into HBM through the double-buffered
// ingest (the dispatcher's prefill and double buffering); VDIF and Mark5B
// frames are validated and header-stripped on the device by the frame
// placement kernel (Mark5B samples recoded to this library's 2-bit codes);
// each input keeps a device stream buffer that carries the samples of an
// unfinished spectrum over to the next chunk, so spectra never depend on the
// chunk size.  Per chunk: auto spectra of each input's selected channel
// (TaskCoreIPP.cpp:437-451), the cross spectrum of the two (TaskCoreIPP.cpp:
// 456-487), phase-cal per spectrum (TaskCoreIPP.cpp:413-418,633-693), written
// to the binary sinks <base>_swspec.bin (fft_ssb_points float32 per spectrum,
// plot_swfile.m:19-29), <base1>_xpol_swspec.bin (fft_ssb_points complex
// float32) and <base>_pcal.bin (tonebins complex float32; FileSink.cpp:68-70
// writes buffers as they are).
#include "fx_internal.h"

#include "spectrometer_fx.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

struct Input {
    fx_ingest* g = nullptr;
    FILE* sink = nullptr;
    FILE* pcal_sink = nullptr;
    int chan = 0;
    void* buf[2] = {nullptr, nullptr};   // stream buffers (carried samples + new chunk)
    int cur = 0;
    size_t cap = 0;                      // bytes per stream buffer
    size_t have = 0;                     // bytes held in buf[cur]
    void* place = nullptr;               // VDIF payload staging
    size_t place_cap = 0;
    uint32_t* valid = nullptr;
    uint64_t* counts = nullptr;
    int* map = nullptr;
    void* fold = nullptr;
    size_t fold_cap = 0;
    uint8_t* wv[2] = {nullptr, nullptr};  // framed input: per packed word, 0 = blanked
    bool eof = false;
};

struct App {
    fx_ctx* ctx = nullptr;
    Input in[2];
    int ninp = 0;
    FILE* xsink = nullptr;
    void* d_auto = nullptr;
    size_t auto_cap = 0;
    void* d_x = nullptr;
    size_t x_cap = 0;
    int* d_any = nullptr;
};

int grow(fx_ctx* ctx, void*& p, size_t& cap, size_t bytes)
{
    if (cap >= bytes) return FX_OK;
    if (p) fx_free(ctx, p);
    p = nullptr;
    cap = 0;
    int rc = fx_malloc(ctx, bytes, &p);
    if (rc == FX_OK) cap = bytes;
    return rc;
}

void close_app(App& a)
{
    for (Input& i : a.in) {
        if (i.g) fx_ingest_close(i.g);
        if (i.sink) fclose(i.sink);
        if (i.pcal_sink) fclose(i.pcal_sink);
        if (a.ctx) {
            for (void* p : {i.buf[0], i.buf[1], i.place, (void*)i.valid, (void*)i.counts, (void*)i.map, i.fold,
                            (void*)i.wv[0], (void*)i.wv[1]})
                if (p) fx_free(a.ctx, p);
        }
    }
    if (a.xsink) fclose(a.xsink);
    if (a.ctx) {
        if (a.d_auto) fx_free(a.ctx, a.d_auto);
        if (a.d_x) fx_free(a.ctx, a.d_x);
        if (a.d_any) fx_free(a.ctx, a.d_any);
        fx_ctx_destroy(a.ctx);
    }
}

FILE* open_sink(const char* base, const char* suffix)
{
    std::string name = std::string(base) + suffix;
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) fx_set_error("fx_swspectrometer_run: cannot open %s", name.c_str());
    return f;
}

// Next chunk of input i appended to its stream buffer; false at the end.
int pull(App& a, Input& in, const fx_swspec_settings* s, size_t payload_bytes, int fps)
{
    const void* d = nullptr;
    size_t n = 0;
    int rc = fx_ingest_next(in.g, &d, &n);
    if (rc) return rc;
    if (n == 0) {
        in.eof = true;
        return FX_OK;
    }
    const void* src = d;
    size_t add = n;
    if (s->format != FX_FMT_RAW) {
        const int fb = s->framebytes, hb = s->headerbytes;
        const size_t nfr = n / (size_t)fb;
        FX_REQUIRE(nfr > 0 && n % (size_t)fb == 0, FX_ERR_ARG, "fx_swspectrometer_run: partial frame");
        // slots counted from the chunk's first frame (frames arrive in order)
        unsigned char hdr[16];
        if ((rc = fx_memcpy_d2h(a.ctx, hdr, d, 16))) return rc;
        long sec0, frame0;
        if (s->format == FX_FMT_VDIF) {
            fx_vdif_header h;
            fx_vdif_decode(hdr, &h);
            sec0 = (long)h.seconds;
            frame0 = (long)h.frame;
        } else {
            fx_mark5b_header h;
            fx_mark5b_decode(hdr, &h);
            sec0 = (long)h.day * 86400L + (long)h.second_of_day;
            frame0 = (long)h.frame;
        }
        add = nfr * payload_bytes;
        if ((rc = grow(a.ctx, in.place, in.place_cap, add))) return rc;
        rc = fx_frames_place(a.ctx, s->format, d, nfr, fb, hb, in.map, 1024, 1, sec0, frame0, fps, (long)nfr,
                             (uint32_t*)in.place, add / 4, in.valid, in.counts, FX_PLACE_CLEAR);
        if (rc) return rc;
        src = in.place;
    }
    FX_REQUIRE(in.have + add <= in.cap, FX_ERR_ARG, "fx_swspectrometer_run: chunk larger than the stream buffer");
    if ((rc = fx_memcpy_d2d(a.ctx, (char*)in.buf[in.cur] + in.have, src, add))) return rc;
    if (s->format != FX_FMT_RAW &&
        fxk_valid_words(fx_ctx_hip_stream(a.ctx), in.valid, (long)(n / (size_t)s->framebytes),
                        (int)(payload_bytes / 4), in.wv[in.cur] + in.have / 4) != FX_OK) {
        fx_set_error("fx_swspectrometer_run: validity launch failed");
        return FX_ERR_HIP;
    }
    in.have += add;
    return FX_OK;
}

} // namespace

extern "C" int fx_swspectrometer_run(const fx_swspec_settings* s, const char* input1, const char* input2,
                                     fx_swspec_stats* stats)
{
    FX_REQUIRE(s && input1 && s->basefilename[0], FX_ERR_ARG, "fx_swspectrometer_run: null argument");
    const fx_spectrometer_settings& c = s->core;
    const int N = c.fft_points, ov = c.fft_overlap_factor, K = c.core_overlapped_ffts, nch = c.source_channels;
    FX_REQUIRE((N & (N - 1)) == 0 && N >= 256 && N <= 32768 && ov >= 1 && N % ov == 0 && (N / ov) % 16 == 0 &&
                   K >= 1,
               FX_ERR_ARG, "fx_swspectrometer_run: fft_points %d, overlap %d, ffts %d", N, ov, K);
    FX_REQUIRE(c.bits_per_sample == 2, FX_ERR_UNSUPPORTED, "fx_swspectrometer_run: %d-bit input", c.bits_per_sample);
    FX_REQUIRE(nch == 1 || nch == 2 || nch == 4 || nch == 8 || nch == 16, FX_ERR_ARG,
               "fx_swspectrometer_run: source_channels %d", nch);
    FX_REQUIRE(s->format == FX_FMT_RAW || s->format == FX_FMT_VDIF || s->format == FX_FMT_MARK5B, FX_ERR_ARG,
               "fx_swspectrometer_run: format %d", s->format);
    FX_REQUIRE(s->spectra_per_chunk >= 1, FX_ERR_ARG, "fx_swspectrometer_run: spectra_per_chunk");
    const int ninp = input2 ? 2 : 1;
    for (int i = 0; i < ninp; i++)
        FX_REQUIRE(s->use_channel_file[i] >= 1 && s->use_channel_file[i] <= nch, FX_ERR_ARG,
                   "fx_swspectrometer_run: use_channel_file[%d] = %d", i, s->use_channel_file[i]);
    const bool xpol = s->calc_xpol && ninp == 2;
    // each input writes its own sinks (swspectrometer.cpp:250-254)
    FX_REQUIRE(ninp == 1 || (s->basefilename[1] && strcmp(s->basefilename[0], s->basefilename[1]) != 0), FX_ERR_ARG,
               "fx_swspectrometer_run: two inputs need two different sink names");
    // phase-cal as the settings derive it (swspectrometer.cpp:178-196)
    int T = 0, R = 0;
    if (c.extract_pcal) {
        FX_REQUIRE(nch == 1, FX_ERR_UNSUPPORTED, "fx_swspectrometer_run: phase-cal on multi-channel sources");
        FX_REQUIRE(c.samplingfreq > 0 && c.pcalharmonicshz > 0 && c.pcaloffsethz >= 0, FX_ERR_ARG,
                   "fx_swspectrometer_run: phase-cal frequencies");
        const long long fs = (long long)c.samplingfreq, off = (long long)c.pcaloffsethz;
        T = (int)(c.samplingfreq / c.pcalharmonicshz);
        R = off ? (int)(fs / std::gcd(fs, off)) : T;
        FX_REQUIRE(T >= 1 && R % T == 0, FX_ERR_UNSUPPORTED,
                   "fx_swspectrometer_run: phase-cal rotator length %d is not a multiple of %d tone bins", R, T);
        if (N % R != 0) T = R = 0;       // extraction switched off (swspectrometer.cpp:185-190)
    }
    const double dphi = 2.0 * M_PI * (-c.pcaloffsethz / (c.samplingfreq > 0 ? c.samplingfreq : 1.0));

    const long L = (long)(K - 1) * (N / ov) + N;          // samples per spectrum (TaskCoreIPP.cpp:495-510)
    const size_t spec_bytes = (size_t)L * (size_t)nch / 4;  // 2-bit samples, nch channels
    const int C1 = N / 2 + 1;
    size_t chunk_bytes = spec_bytes * (size_t)s->spectra_per_chunk;
    size_t payload_bytes = 0;
    int fps = 1;
    if (s->format != FX_FMT_RAW) {
        FX_REQUIRE(s->framebytes > s->headerbytes && (s->headerbytes == 16 || s->headerbytes == 32) &&
                       s->framebytes % 8 == 0 && s->frames_per_second > 0,
                   FX_ERR_ARG, "fx_swspectrometer_run: frames of %d/%d bytes, %d per second", s->framebytes,
                   s->headerbytes, s->frames_per_second);
        payload_bytes = (size_t)(s->framebytes - s->headerbytes);
        fps = s->frames_per_second;
        const size_t nfr = (chunk_bytes + payload_bytes - 1) / payload_bytes;
        chunk_bytes = nfr * (size_t)s->framebytes;
    }
    const size_t new_max = s->format != FX_FMT_RAW ? chunk_bytes / (size_t)s->framebytes * payload_bytes : chunk_bytes;

    App a;
    a.ninp = ninp;
    int rc = fx_ctx_create(0, &a.ctx);
    if (rc) return rc;
    const char* names[2] = {input1, input2};
    for (int i = 0; i < ninp && rc == FX_OK; i++) {
        Input& in = a.in[i];
        in.chan = s->use_channel_file[i] - 1;
        in.cap = spec_bytes + new_max + 16;
        size_t c0 = 0, c1 = 0;
        rc = grow(a.ctx, in.buf[0], c0, in.cap);
        if (!rc) rc = grow(a.ctx, in.buf[1], c1, in.cap);
        if (!rc && s->format != FX_FMT_RAW) {
            void *v = nullptr, *cn = nullptr, *m = nullptr;
            const size_t nfr = chunk_bytes / (size_t)s->framebytes;
            rc = fx_malloc(a.ctx, 8 * nfr, &v);
            if (!rc) rc = fx_malloc(a.ctx, 32, &cn);
            if (!rc) rc = fx_malloc(a.ctx, 1024 * sizeof(int), &m);
            in.valid = (uint32_t*)v;
            in.counts = (uint64_t*)cn;
            in.map = (int*)m;
            if (!rc) {
                std::vector<int> map(1024, 0);       // every thread of the file is this input
                rc = fx_memcpy_h2d(a.ctx, in.map, map.data(), map.size() * sizeof(int));
            }
            void *w0 = nullptr, *w1 = nullptr;
            if (!rc) rc = fx_malloc(a.ctx, in.cap / 4 + 16, &w0);
            if (!rc) rc = fx_malloc(a.ctx, in.cap / 4 + 16, &w1);
            in.wv[0] = (uint8_t*)w0;
            in.wv[1] = (uint8_t*)w1;
            if (!rc && !a.d_any) {
                void* p = nullptr;
                rc = fx_malloc(a.ctx, 16, &p);
                a.d_any = (int*)p;
            }
        }
        const char* base = s->basefilename[i];
        if (!rc && !(in.sink = open_sink(base, "_swspec.bin"))) rc = FX_ERR_ARG;
        if (!rc && T > 0 && !(in.pcal_sink = open_sink(base, "_pcal.bin"))) rc = FX_ERR_ARG;
        if (!rc) rc = fx_ingest_open(a.ctx, names[i], 0, chunk_bytes, 2, &in.g);
    }
    if (!rc && xpol && !(a.xsink = open_sink(s->basefilename[0], "_xpol_swspec.bin"))) rc = FX_ERR_ARG;
    if (rc) {
        close_app(a);
        return rc;
    }

    const auto t0 = std::chrono::steady_clock::now();
    long total = 0, chunks = 0;
    std::vector<float> h_auto, h_x;
    std::vector<fx_cf64> pc;
    std::vector<float> pc32;
    for (;;) {
        bool eof = false;
        for (int i = 0; i < ninp && !rc; i++) {
            if (!a.in[i].eof) rc = pull(a, a.in[i], s, payload_bytes, fps);
            eof = eof || a.in[i].eof;
        }
        if (rc) break;
        // whole spectra held by every input (the inputs run in lockstep)
        long ns = -1;
        for (int i = 0; i < ninp; i++) {
            const long n_i = (long)(a.in[i].have / spec_bytes);
            ns = ns < 0 ? n_i : std::min(ns, n_i);
        }
        if (ns > 0) {
            const size_t ntime = (size_t)ns * (size_t)L;
            if ((rc = grow(a.ctx, a.d_auto, a.auto_cap, (size_t)ns * nch * C1 * sizeof(float)))) break;
            h_auto.resize((size_t)ns * nch * C1);
            for (int i = 0; i < ninp && !rc; i++) {
                Input& in = a.in[i];
                const uint32_t* w = (const uint32_t*)in.buf[in.cur];
                // a framed input whose processed words include blanked ones
                // goes through the masked frame path (blanked words as zeros)
                const uint8_t* wv = in.wv[in.cur];
                bool blanked = false;
                if (wv) {
                    int any = 0;
                    rc = fx_memset(a.ctx, a.d_any, 0, sizeof(int));
                    if (!rc && fxk_any_zero(fx_ctx_hip_stream(a.ctx), wv, ntime * (size_t)nch / 16, a.d_any) != FX_OK)
                        rc = FX_ERR_HIP;
                    if (!rc) rc = fx_memcpy_d2h(a.ctx, &any, a.d_any, sizeof(int));
                    blanked = any != 0;
                }
                if (!rc && blanked) {
                    rc = grow(a.ctx, a.d_x, a.x_cap, (size_t)ns * C1 * sizeof(fx_cf32));
                    if (!rc)
                        rc = fx_spec_xpol_impl(a.ctx, w, wv, nch, in.chan, w, wv, nch, in.chan, ntime, N, ov, K,
                                               c.window, 0, (fx_cf32*)a.d_x, nullptr);
                    h_x.resize((size_t)ns * C1 * 2);
                    if (!rc) rc = fx_memcpy_d2h(a.ctx, h_x.data(), a.d_x, h_x.size() * sizeof(float));
                    std::vector<float> re((size_t)C1);
                    for (long sp = 0; sp < ns && !rc; sp++) {
                        for (int k = 0; k < C1; k++) re[(size_t)k] = h_x[2 * ((size_t)sp * C1 + k)];
                        if (fwrite(re.data(), sizeof(float), C1, in.sink) != (size_t)C1) rc = FX_ERR_STATE;
                    }
                } else if (!rc) {
                    long got = 0;
                    rc = fx_spectrometer_ex(a.ctx, w, ntime, nch, N, ov, K, c.window, 0, 1, (float*)a.d_auto, &got);
                    if (!rc) rc = fx_memcpy_d2h(a.ctx, h_auto.data(), a.d_auto, h_auto.size() * sizeof(float));
                    for (long sp = 0; sp < ns && !rc; sp++)
                        if (fwrite(&h_auto[((size_t)sp * nch + in.chan) * C1], sizeof(float), C1, in.sink) !=
                            (size_t)C1)
                            rc = FX_ERR_STATE;
                }
                if (!rc && T > 0) {
                    rc = fx_spec_pcal(a.ctx, w, ns, L, total, N, ov, K, R, T, dphi, in.fold, in.fold_cap, pc,
                                      blanked ? wv : nullptr);
                    pc32.resize(pc.size() * 2);
                    for (size_t k = 0; k < pc.size(); k++) {
                        pc32[2 * k] = (float)pc[k].re;
                        pc32[2 * k + 1] = (float)pc[k].im;
                    }
                    if (!rc && fwrite(pc32.data(), sizeof(float), pc32.size(), in.pcal_sink) != pc32.size())
                        rc = FX_ERR_STATE;
                }
            }
            if (!rc && xpol) {
                rc = grow(a.ctx, a.d_x, a.x_cap, (size_t)ns * C1 * sizeof(fx_cf32));
                if (!rc)
                    rc = fx_spec_xpol_impl(a.ctx, (const uint32_t*)a.in[0].buf[a.in[0].cur], a.in[0].wv[a.in[0].cur],
                                           nch, a.in[0].chan, (const uint32_t*)a.in[1].buf[a.in[1].cur],
                                           a.in[1].wv[a.in[1].cur], nch, a.in[1].chan, ntime, N, ov, K, c.window,
                                           0, (fx_cf32*)a.d_x, nullptr);
                h_x.resize((size_t)ns * C1 * 2);
                if (!rc) rc = fx_memcpy_d2h(a.ctx, h_x.data(), a.d_x, h_x.size() * sizeof(float));
                if (!rc && fwrite(h_x.data(), sizeof(float), h_x.size(), a.xsink) != h_x.size()) rc = FX_ERR_STATE;
            }
            if (rc) break;
            total += ns;
            // carry the unfinished spectrum's samples to the other buffer
            for (int i = 0; i < ninp && !rc; i++) {
                Input& in = a.in[i];
                const size_t used = (size_t)ns * spec_bytes, left = in.have - used;
                if (left) rc = fx_memcpy_d2d(a.ctx, in.buf[1 - in.cur], (char*)in.buf[in.cur] + used, left);
                if (left && !rc && in.wv[0])
                    rc = fx_memcpy_d2d(a.ctx, in.wv[1 - in.cur], in.wv[in.cur] + used / 4, left / 4);
                in.cur = 1 - in.cur;
                in.have = left;
            }
            if (rc) break;
        }
        chunks++;
        // the run ends when any input ends (TaskDispatcher.cpp:208-218)
        if (eof) break;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        stats->spectra = total;
        stats->chunks = chunks;
        stats->seconds = secs;
        // signal time of the spectra over wall time (TaskDispatcher.cpp:183-192):
        // a spectrum spans L samples
        stats->realtime_factor =
            c.samplingfreq > 0 && secs > 0 ? (double)total * (double)L / c.samplingfreq / secs : 0.0;
    }
    if (rc == FX_ERR_STATE) fx_set_error("fx_swspectrometer_run: sink write failed");
    close_app(a);
    return rc;
}

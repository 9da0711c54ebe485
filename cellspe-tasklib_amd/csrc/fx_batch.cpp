// fx_batch.cpp -- many subbands with uneven FFT lengths and input counts in
// one launch, and the autotuner that picks each subband's F-stage sub-chunk
// depth and X-engine tile.
//
// Every subband gets its own correlator on its own HIP stream.  With a graph,
// the batch's work is captured once: the caller's stream forks to the subband
// streams (an event each way), so the instantiated graph holds one branch per
// subband and the branches run concurrently; fx_batch_run replays it with one
// hipGraphLaunch.  All clears inside it are fill kernels (no memset nodes).
//
// The autotuner times the F-stage + X-engine of one chunk of a subband shape
// on synthetic 2-bit input for each (sub-chunk depth, X tile) candidate with
// HIP events and caches the fastest per (nfft, ninp, precision, chunk) in a
// text file.
#include "fx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Subband {
    fx_subband_desc d{};
    fx_ctx* ctx = nullptr;
    fx_corr* corr = nullptr;
    fx_cf32* vis = nullptr;
    fx_pcal* pcal = nullptr;   // all inputs of the subband, one fold + one final launch
    long chunk = 0, sub = 0;
    int xtile = 128;
    // persistent modes: the whole integration's channelized samples, the
    // device phase models, channels per X item and F groups per item
    uint32_t* chan = nullptr;
    fx_phase_model* pm_dev = nullptr;
    int cb = 1, gpi = 4;
};

long round8(long n) { return (n + 7) / 8 * 8; }

long default_chunk(const fx_subband_desc& d) { return std::min<long>(4096, round8(std::max<long>(d.nspectra, 8))); }

bool lookup(const char* file, int nfft, int ninp, int prec, long chunk, long* sub, int* xtile)
{
    if (!file) return false;
    FILE* f = fopen(file, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        int n, i, p, x;
        long ch, s;
        double ms;
        if (line[0] == '#') continue;
        if (sscanf(line, "%d %d %d %ld %ld %d %lf", &n, &i, &p, &ch, &s, &x, &ms) == 7 && n == nfft &&
            i == ninp && p == prec && ch == chunk) {
            *sub = s;
            *xtile = x;
            found = true;   // the last entry wins
        }
    }
    fclose(f);
    return found;
}

// enqueue one subband's work on its own stream
int enqueue(Subband& b)
{
    const fx_subband_desc& d = b.d;
    int rc = fx_corr_reset(b.corr);
    if (!rc) rc = fx_corr_process(b.corr, d.d_words, d.stride_words, 0, d.nspectra);
    if (!rc) rc = fx_corr_dump(b.corr, b.vis);
    if (!rc && b.pcal) {
        rc = fx_pcal_clear(b.pcal);
        if (!rc) rc = fx_pcal_adjust_sample_offset(b.pcal, 0);
        if (!rc) rc = fx_pcal_extract_2bit_inputs(b.pcal, d.d_words, d.stride_words, (size_t)d.nspectra * d.nfft);
        if (!rc) rc = fx_pcal_final_async(b.pcal);
    }
    return rc;
}

} // namespace

struct fx_batch {
    fx_ctx* ctx = nullptr;
    int prec = FX_PREC_F16;
    bool use_graph = false;
    std::vector<Subband> sb;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // persistent modes (FX_BATCH_PERSISTENT, FX_BATCH_PHASED)
    int mode = FX_BATCH_EAGER;
    fx_psub* d_subs = nullptr;
    fx_pitem* d_items = nullptr;
    int* d_ctrs = nullptr;
    int nctrs = 0, nitems = 0;
    int phase_end[3] = {0, 0, 0};   // phased: launches over items [0,e0), [e0,e1), [e1,e2)
    int nblocks = 0;
    size_t lds = 0;
};

namespace {

// Work table of the persistent modes.  Counters: FX_PCTR_FIRST + 2 s counts
// the finished F items of subband s, + 1 its finished phase-cal folds.
// fused (FX_BATCH_PERSISTENT): the folds, then the F and X items of the
// subbands interleaved so that a subband's X items follow the next subband's
// F items (its own F items are then mostly done when they are taken), then
// the tone items; phased: folds and F items, X items, tones as three ranges.
int build_table(fx_batch* b, bool fused, std::vector<fx_pitem>& items, int* phase_end)
{
    const int ns = (int)b->sb.size();
    std::vector<int> nF((size_t)ns, 0), nP((size_t)ns, 0);
    auto fctr = [](int s) { return FX_PCTR_FIRST + 2 * s; };
    auto pctr = [](int s) { return FX_PCTR_FIRST + 2 * s + 1; };
    auto item = [](int type, int sb, long a, long bb, int c, int xt, int wc, int wv, int dc) {
        fx_pitem it;
        it.type = type;
        it.sb = sb;
        it.a = a;
        it.b = bb;
        it.c = c;
        it.xtile = xt;
        it.wait_ctr = wc;
        it.wait_val = wv;
        it.done_ctr = dc;
        it.pad = 0;
        return it;
    };
    auto pcal_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (!r.pcal) return;
        fx_psub S{};
        fx_pcal_device_state(r.pcal, &S, nullptr);
        const long nbin = S.nfold / 16;
        const long wpi = nbin * std::max<long>(1, 65536 / nbin);   // about 1M samples per fold item
        const long nw = r.d.nspectra * r.d.nfft / 16;
        for (int y = 0; y < r.d.ninp; y++)
            for (long w0 = 0; w0 < nw; w0 += wpi) {
                items.push_back(item(FX_PI_PCAL, s, y, w0, (int)wpi, 0, -1, 0, pctr(s)));
                nP[(size_t)s]++;
            }
    };
    auto f_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        const long F = fx_persist_group_items(r.d.nfft);
        const long G = (r.d.nspectra * r.d.ninp + F - 1) / F;
        for (long g0 = 0; g0 < G; g0 += r.gpi) {
            items.push_back(item(FX_PI_F, s, g0, std::min<long>(r.gpi, G - g0), 0, 0, -1, 0, fctr(s)));
            nF[(size_t)s]++;
        }
    };
    auto x_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        const int nb = (r.d.ninp + r.xtile - 1) / r.xtile;
        const long C = r.d.nfft / 2;
        for (int t = 0; t < nb * (nb + 1) / 2; t++)
            for (long c0 = 0; c0 < C; c0 += r.cb)
                items.push_back(
                    item(FX_PI_X, s, c0, std::min<long>(r.cb, C - c0), t, r.xtile, fctr(s), nF[(size_t)s], -1));
    };
    auto t_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (!r.pcal) return;
        for (int y = 0; y < r.d.ninp; y++) items.push_back(item(FX_PI_TONE, s, y, 0, 0, 0, pctr(s), nP[(size_t)s], -1));
    };
    // the largest subbands first (their X items take longest)
    std::vector<int> order((size_t)ns);
    for (int s = 0; s < ns; s++) order[(size_t)s] = s;
    auto work = [&](int s) {
        const fx_subband_desc& d = b->sb[(size_t)s].d;
        return (double)d.ninp * d.ninp * (double)d.nspectra * d.nfft;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return work(a) > work(c); });
    items.clear();
    for (int s : order) pcal_items(s);
    if (fused) {
        for (int k = 0; k < ns; k++) {
            f_items(order[(size_t)k]);
            if (k) x_items(order[(size_t)k - 1]);
        }
        x_items(order[(size_t)ns - 1]);
        phase_end[0] = phase_end[1] = (int)items.size();
    } else {
        for (int s : order) f_items(s);
        phase_end[0] = (int)items.size();
        for (int s : order) x_items(s);
        phase_end[1] = (int)items.size();
    }
    for (int s : order) t_items(s);
    phase_end[2] = (int)items.size();
    // every wait names items that come earlier in the table: the workgroups
    // that took them are running when the wait starts, so it ends
    std::vector<int> seen((size_t)(FX_PCTR_FIRST + 2 * ns), 0);
    for (const fx_pitem& it : items) {
        if (it.wait_ctr >= 0 && seen[(size_t)it.wait_ctr] != it.wait_val) {
            fx_set_error("fx_batch_create: work table waits for items that do not precede it");
            return FX_ERR_STATE;
        }
        if (it.done_ctr >= 0) seen[(size_t)it.done_ctr]++;
    }
    return FX_OK;
}

// the persistent run's launches on stream s: clear the counters and folds, then
// one fused launch or three phased ones
int enqueue_persist(fx_batch* b, hipStream_t s)
{
    if (fxk_fill32(s, b->d_ctrs, 0u, (size_t)b->nctrs) != FX_OK) return FX_ERR_HIP;
    for (const Subband& r : b->sb) {
        if (!r.pcal) continue;
        fx_psub S{};
        fx_pcal_device_state(r.pcal, &S, nullptr);
        if (fxk_fill32(s, S.pacc, 0u, (size_t)S.nfold * r.d.ninp * 2) != FX_OK) return FX_ERR_HIP;
    }
    const float2* tw = fx_ctx_twiddles(b->ctx);
    if (b->mode == FX_BATCH_PERSISTENT)
        return fxk_batch_persistent(s, tw, b->d_subs, b->d_items, 0, b->nitems, b->d_ctrs, 0, b->nblocks, 1, b->lds);
    int i0 = 0;
    for (int ph = 0; ph < 3; ph++) {
        int rc = fxk_batch_persistent(s, tw, b->d_subs, b->d_items, i0, b->phase_end[ph], b->d_ctrs, ph, b->nblocks, 0,
                                      b->lds);
        if (rc) return rc;
        i0 = b->phase_end[ph];
    }
    return FX_OK;
}

bool lookup_persist(const char* file, int nfft, int ninp, long nspectra, int* xtile, int* cb, int* gpi)
{
    if (!file) return false;
    FILE* f = fopen(file, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        int n, i, x, c, g;
        long t;
        double ms;
        if (sscanf(line, "P %d %d %ld %d %d %d %lf", &n, &i, &t, &x, &c, &g, &ms) == 7 && n == nfft && i == ninp &&
            t == nspectra) {
            *xtile = x;
            *cb = c;
            *gpi = g;
            found = true;
        }
    }
    fclose(f);
    return found;
}

// defaults: 64-input tiles, X items of about 2048 spectra of channels, 4 F groups per item
void persist_defaults(const fx_subband_desc& d, int* xtile, int* cb, int* gpi)
{
    *xtile = 64;
    *cb = (int)std::max<long>(1, std::min<long>(d.nfft / 2, 2048 / std::max<long>(1, d.nspectra)));
    *gpi = 4;
}

// table of one subband only (the autotuner's timing runs)
int build_table_one(fx_batch* b, int s, std::vector<fx_pitem>& items, int* phase_end)
{
    fx_batch one;
    one.sb.push_back(b->sb[(size_t)s]);
    int rc = build_table(&one, false, items, phase_end);
    for (fx_pitem& it : items) {
        it.sb = s;
        if (it.done_ctr >= 0) it.done_ctr += 2 * s;
        if (it.wait_ctr >= 0) it.wait_ctr += 2 * s;
    }
    one.sb.clear();   // owned by b
    return rc;
}

// Time one subband's F items (ph = 0) or X items (ph = 1) with its current
// tuning, phased, in ms (the second of two runs).
int time_phase(fx_batch* b, int s, int ph, float* ms)
{
    std::vector<fx_pitem> items;
    int pe[3];
    int rc = build_table_one(b, s, items, pe);
    if (rc) return rc;
    fx_pitem* d = nullptr;
    rc = fx_malloc(b->ctx, items.size() * sizeof(fx_pitem), (void**)&d);
    if (rc) return rc;
    rc = fx_memcpy_h2d(b->ctx, d, items.data(), items.size() * sizeof(fx_pitem));
    void *e0 = nullptr, *e1 = nullptr;
    if (!rc) rc = fx_event_create(b->ctx, &e0);
    if (!rc) rc = fx_event_create(b->ctx, &e1);
    hipStream_t st = fx_ctx_hip_stream(b->ctx);
    const float2* tw = fx_ctx_twiddles(b->ctx);
    const int i0 = ph == 0 ? 0 : pe[0], i1 = ph == 0 ? pe[0] : pe[1];
    for (int rep = 0; rep < 2 && !rc; rep++) {
        rc = fxk_fill32(st, b->d_ctrs, 0u, (size_t)b->nctrs);
        if (!rc) rc = fx_event_record(b->ctx, e0);
        if (!rc) rc = fxk_batch_persistent(st, tw, b->d_subs, d, i0, i1, b->d_ctrs, ph, b->nblocks, 0, b->lds);
        if (!rc) rc = fx_event_record(b->ctx, e1);
        if (!rc) rc = fx_event_elapsed_ms(e0, e1, ms);
    }
    if (e0) fx_event_destroy(e0);
    if (e1) fx_event_destroy(e1);
    fx_ctx_sync(b->ctx);
    fx_free(b->ctx, d);
    return rc;
}

// Autotuner of the persistent modes for subband s (every later subband of the
// same shape takes its result): F groups per item from {1, 2, 4, 8, 16}, then
// channels per X item (64 x 64-input tiles) from {1, 2, ..., 64},
// each by timing that subband's items alone on the device (its real input);
// the choice is appended to tune_file as "P nfft ninp nspectra xtile cb gpi ms".
int autotune_persist(fx_batch* b, int s, const char* tune_file)
{
    Subband& r = b->sb[(size_t)s];
    float best = 1e30f, ms = 0.f, fms = 0.f;
    int best_g = r.gpi;
    for (int g : {1, 2, 4, 8, 16}) {
        r.gpi = g;
        int rc = time_phase(b, s, 0, &ms);
        if (rc) return rc;
        if (ms < best) {
            best = ms;
            best_g = g;
        }
    }
    r.gpi = best_g;
    fms = best;
    best = 1e30f;
    int best_x = r.xtile, best_c = r.cb;
    const long C = r.d.nfft / 2;
    for (int xt : {64}) {   // the persistent kernel's X items are 64 x 64 input tiles
        for (int cb = 1; cb <= 64 && cb <= C; cb *= 2) {
            r.xtile = xt;
            r.cb = cb;
            int rc = time_phase(b, s, 1, &ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                best_x = xt;
                best_c = cb;
            }
        }
    }
    r.xtile = best_x;
    r.cb = best_c;
    if (tune_file) {
        FILE* f = fopen(tune_file, "a");
        if (f) {
            fprintf(f, "P %d %d %ld %d %d %d %.4f\n", r.d.nfft, r.d.ninp, r.d.nspectra, r.xtile, r.cb, r.gpi,
                    fms + best);
            fclose(f);
        }
    }
    return FX_OK;
}

int batch_create_persist(fx_ctx* ctx, const fx_subband_desc* sb, int nsub, int prec, int mode, const char* tune_file,
                         fx_batch** out)
{
    FX_REQUIRE(prec == FX_PREC_F16, FX_ERR_UNSUPPORTED,
               "fx_batch_create: the persistent modes run the FP16 engine (precision %d)", prec);
    fx_batch* b = new fx_batch;
    b->ctx = ctx;
    b->prec = prec;
    b->mode = mode;
    b->use_graph = true;
    b->sb.resize((size_t)nsub);
    int rc = FX_OK;
    size_t lds = fx_persist_lds(FX_PI_TONE, 0);
    std::vector<fx_psub> hs((size_t)nsub);
    std::vector<bool> tuned((size_t)nsub, false);
    for (int s = 0; s < nsub && !rc; s++) {
        Subband& r = b->sb[(size_t)s];
        r.d = sb[s];
        const fx_subband_desc& d = r.d;
        const bool pow2 = d.nfft >= 256 && d.nfft <= 16384 && (d.nfft & (d.nfft - 1)) == 0;
        if (!pow2 || d.ninp < 1 || d.nspectra < 1 || !d.d_words || d.stride_words % 2 ||
            ((uintptr_t)d.d_words % 8) || d.stride_words < (size_t)d.nspectra * d.nfft / 16) {
            fx_set_error("fx_batch_create: subband %d: nfft %d, %d inputs, %ld spectra, word stride %zu", s, d.nfft,
                         d.ninp, d.nspectra, d.stride_words);
            rc = FX_ERR_ARG;
            break;
        }
        r.chunk = d.nspectra;
        persist_defaults(d, &r.xtile, &r.cb, &r.gpi);
        if (tune_file) tuned[(size_t)s] = lookup_persist(tune_file, d.nfft, d.ninp, d.nspectra, &r.xtile, &r.cb, &r.gpi);
        const size_t C = (size_t)d.nfft / 2, P = (size_t)d.ninp * (d.ninp + 1) / 2;
        rc = fx_malloc(ctx, C * (size_t)d.nspectra * (size_t)d.ninp * 4, (void**)&r.chan);
        if (!rc) rc = fx_malloc(ctx, P * C * sizeof(fx_cf32), (void**)&r.vis);
        if (!rc && d.pm) {
            rc = fx_malloc(ctx, sizeof(fx_phase_model) * (size_t)d.ninp, (void**)&r.pm_dev);
            if (!rc) rc = fx_memcpy_h2d(ctx, r.pm_dev, d.pm, sizeof(fx_phase_model) * (size_t)d.ninp);
        }
        fx_psub& S = hs[(size_t)s];
        S.words = d.d_words;
        S.stride = d.stride_words;
        S.nfft = d.nfft;
        S.ninp = d.ninp;
        S.nt = d.nspectra;
        S.pm = r.pm_dev;
        // a 2-bit stream through the level table has E[x^2] ~ 4.316: the FFT
        // output's sigma sqrt(N * 4.316) maps to 1 before the FP16 cast
        S.sc = (float)(1.0 / sqrt((double)d.nfft * 4.316));
        S.chan = r.chan;
        S.vis = (float2*)r.vis;
        if (!rc && d.pcal) {
            rc = fx_pcal_create_inputs(ctx, d.bandwidth_hz, d.pcal_spacing_hz, d.pcal_offset_hz, 0, d.ninp, &r.pcal);
            if (!rc) rc = fx_pcal_device_state(r.pcal, &S, nullptr);
            if (!rc && (S.nfold % 16 || S.nfold > 4096)) {
                fx_set_error("fx_batch_create: subband %d: a phase-cal fold of %d bins (the persistent modes fold "
                             "multiples of 16 up to 4096)", s, S.nfold);
                rc = FX_ERR_UNSUPPORTED;
            }
            lds = std::max(lds, fx_persist_lds(FX_PI_PCAL, S.nfold));
        }
        lds = std::max(lds, std::max(fx_persist_lds(FX_PI_F, d.nfft), fx_persist_lds(FX_PI_X, 128)));
    }
    b->nctrs = FX_PCTR_FIRST + 2 * nsub;
    if (!rc) rc = fx_malloc(ctx, sizeof(fx_psub) * (size_t)nsub, (void**)&b->d_subs);
    if (!rc) rc = fx_memcpy_h2d(ctx, b->d_subs, hs.data(), sizeof(fx_psub) * (size_t)nsub);
    if (!rc) rc = fx_malloc(ctx, sizeof(int) * (size_t)b->nctrs, (void**)&b->d_ctrs);
    if (!rc) {
        b->lds = lds;
        b->nblocks = fxk_batch_persistent_blocks(lds);
        if (b->nblocks < 1) rc = FX_ERR_HIP;   // fx_last_error() says why
    }
    // tune each shape once (the tune file already holds the others)
    for (int s = 0; s < nsub && !rc; s++) {
        if (!tune_file || tuned[(size_t)s]) continue;
        rc = autotune_persist(b, s, tune_file);
        const fx_subband_desc& d = b->sb[(size_t)s].d;
        for (int q = s + 1; q < nsub; q++) {
            Subband& o = b->sb[(size_t)q];
            if (!tuned[(size_t)q] && o.d.nfft == d.nfft && o.d.ninp == d.ninp && o.d.nspectra == d.nspectra) {
                o.xtile = b->sb[(size_t)s].xtile;
                o.cb = b->sb[(size_t)s].cb;
                o.gpi = b->sb[(size_t)s].gpi;
                tuned[(size_t)q] = true;
            }
        }
    }
    std::vector<fx_pitem> items;
    if (!rc) rc = build_table(b, mode == FX_BATCH_PERSISTENT, items, b->phase_end);
    if (!rc) {
        b->nitems = (int)items.size();
        rc = fx_malloc(ctx, sizeof(fx_pitem) * items.size(), (void**)&b->d_items);
    }
    if (!rc) rc = fx_memcpy_h2d(ctx, b->d_items, items.data(), sizeof(fx_pitem) * items.size());
    if (!rc) {
        // one graph: the clears and the launch(es)
        hipStream_t s0 = fx_ctx_hip_stream(ctx);
        if (hipStreamSynchronize(s0) != hipSuccess ||
            hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = FX_ERR_HIP;
        } else {
            int crc = enqueue_persist(b, s0);
            hipGraph_t g = nullptr;
            const bool ended = hipStreamEndCapture(s0, &g) == hipSuccess;
            if (crc || !ended) {
                if (g) hipGraphDestroy(g);
                rc = crc ? crc : FX_ERR_HIP;
                fx_set_error("fx_batch_create: graph capture of the persistent run failed");
            } else {
                b->graph = g;
                if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                    fx_set_error("fx_batch_create: graph instantiation failed");
                    rc = FX_ERR_HIP;
                }
            }
        }
        hipGetLastError();
    }
    if (rc) {
        fx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return FX_OK;
}

} // namespace

extern "C" int fx_autotune(fx_ctx* ctx, int nfft, int ninp, int prec, long chunk_spectra, const char* tune_file,
                           long* sub_out, int* xtile_out)
{
    FX_REQUIRE(ctx && nfft >= 256 && ninp >= 1, FX_ERR_ARG, "fx_autotune: bad shape");
    FX_REQUIRE(prec == FX_PREC_F32 || prec == FX_PREC_F16 || prec == FX_PREC_FP8, FX_ERR_ARG,
               "fx_autotune: precision %d has no F-stage", prec);
    FX_REQUIRE(chunk_spectra >= 8 && chunk_spectra % 8 == 0, FX_ERR_ARG,
               "fx_autotune: chunk of %ld spectra (a positive multiple of 8)", chunk_spectra);
    // timed on one chunk of the shape that will run
    const long T = chunk_spectra;
    const size_t wpi = (size_t)T * nfft / 16;
    void* words = nullptr;
    int rc = fx_malloc(ctx, (size_t)ninp * wpi * 4, &words);
    if (rc) return rc;
    for (int i = 0; i < ninp && !rc; i++)
        rc = fx_synth_station(ctx, 99, i, 0.05, i % 8, 0.0, 0.0, 0.9816, (size_t)T * nfft,
                              (uint32_t*)words + (size_t)i * wpi);
    fx_corr_config cfg = {ninp, nfft, T, prec, 0};
    fx_corr* c = nullptr;
    if (!rc) rc = fx_corr_create(ctx, &cfg, &c);
    void *e0 = nullptr, *e1 = nullptr;
    if (!rc) rc = fx_event_create(ctx, &e0);
    if (!rc) rc = fx_event_create(ctx, &e1);
    const long subs[] = {64, 256, 1024, 4096};
    const int tiles[] = {128, 64};
    long best_sub = 0;
    int best_tile = 128;
    float best = 1e30f;
    for (long s : subs) {
        for (int x : tiles) {
            if (rc) break;
            if (x == 64 && prec == FX_PREC_FP8) continue;
            if (s > T) continue;
            rc = fx_corr_tune(c, s, x);
            float ms = 0.f, tot = 0.f;
            for (int rep = 0; rep < 3 && !rc; rep++) {
                rc = fx_event_record(ctx, e0);
                if (!rc) rc = fx_corr_process(c, (const uint32_t*)words, wpi, 0, T);
                if (!rc) rc = fx_event_record(ctx, e1);
                if (!rc) rc = fx_event_elapsed_ms(e0, e1, &ms);
                if (rep > 0) tot += ms;   // first run warms up
            }
            if (!rc && tot < best) {
                best = tot;
                best_sub = s;
                best_tile = x;
            }
        }
    }
    if (e0) fx_event_destroy(e0);
    if (e1) fx_event_destroy(e1);
    if (c) fx_corr_destroy(c);
    fx_free(ctx, words);
    if (rc) return rc;
    if (tune_file) {
        FILE* f = fopen(tune_file, "a");
        if (f) {
            fprintf(f, "%d %d %d %ld %ld %d %.4f\n", nfft, ninp, prec, T, best_sub, best_tile, best / 2.0f);
            fclose(f);
        }
    }
    if (sub_out) *sub_out = best_sub;
    if (xtile_out) *xtile_out = best_tile;
    return FX_OK;
}

extern "C" int fx_batch_destroy(fx_batch* b)
{
    if (!b) return FX_OK;
    if (b->ctx) hipStreamSynchronize(fx_ctx_hip_stream(b->ctx));
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->fork) hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) hipEventDestroy(e);
    for (Subband& s : b->sb) {
        if (s.pcal) fx_pcal_destroy(s.pcal);
        if (s.corr) fx_corr_destroy(s.corr);
        fx_ctx* own = s.ctx ? s.ctx : b->ctx;
        if (s.vis) fx_free(own, s.vis);
        if (s.chan) fx_free(own, s.chan);
        if (s.pm_dev) fx_free(own, s.pm_dev);
        if (s.ctx) fx_ctx_destroy(s.ctx);
    }
    if (b->d_subs) fx_free(b->ctx, b->d_subs);
    if (b->d_items) fx_free(b->ctx, b->d_items);
    if (b->d_ctrs) fx_free(b->ctx, b->d_ctrs);
    delete b;
    return FX_OK;
}

extern "C" int fx_batch_create(fx_ctx* ctx, const fx_subband_desc* sb, int nsub, int prec, int use_graph,
                               const char* tune_file, fx_batch** out)
{
    FX_REQUIRE(ctx && sb && out && nsub >= 1, FX_ERR_ARG, "fx_batch_create: bad arguments");
    FX_REQUIRE(prec == FX_PREC_F32 || prec == FX_PREC_F16 || prec == FX_PREC_FP8, FX_ERR_ARG,
               "fx_batch_create: precision %d (FP32, FP16 or FP8 F->X)", prec);
    *out = nullptr;
    FX_REQUIRE(use_graph >= FX_BATCH_EAGER && use_graph <= FX_BATCH_PHASED, FX_ERR_ARG,
               "fx_batch_create: mode %d", use_graph);
    if (use_graph >= FX_BATCH_PERSISTENT) return batch_create_persist(ctx, sb, nsub, prec, use_graph, tune_file, out);
    fx_batch* b = new fx_batch;
    b->ctx = ctx;
    b->prec = prec;
    b->use_graph = use_graph != 0;
    b->mode = use_graph;
    b->sb.resize((size_t)nsub);
    int rc = FX_OK;
    int device = 0;
    hipGetDevice(&device);
    for (int s = 0; s < nsub && !rc; s++) {
        Subband& r = b->sb[(size_t)s];
        r.d = sb[s];
        const fx_subband_desc& d = r.d;
        if (d.nfft < 256 || d.ninp < 1 || d.nspectra < 8 || d.nspectra % 8 || !d.d_words) {
            fx_set_error("fx_batch_create: subband %d: nfft %d, %d inputs, %ld spectra (a multiple of 8)", s,
                         d.nfft, d.ninp, d.nspectra);
            rc = FX_ERR_ARG;
            break;
        }
        r.chunk = default_chunk(d);
        rc = fx_ctx_create(device, &r.ctx);
        fx_corr_config cfg = {d.ninp, d.nfft, r.chunk, prec, d.pm ? 1 : 0};
        if (!rc) rc = fx_corr_create(r.ctx, &cfg, &r.corr);
        if (!rc && d.pm) rc = fx_corr_set_phase_model(r.corr, d.pm);
        long sub = 0;
        int xtile = 128;
        if (!rc && tune_file && !lookup(tune_file, d.nfft, d.ninp, prec, r.chunk, &sub, &xtile))
            rc = fx_autotune(r.ctx, d.nfft, d.ninp, prec, r.chunk, tune_file, &sub, &xtile);
        if (!rc && tune_file) rc = fx_corr_tune(r.corr, sub, xtile);
        if (!rc) {
            r.sub = fx_corr_fstage_depth(r.corr);
            r.xtile = tune_file ? xtile : 128;
        }
        const size_t P = (size_t)d.ninp * (d.ninp + 1) / 2;
        if (!rc) rc = fx_malloc(r.ctx, P * (size_t)(d.nfft / 2) * sizeof(fx_cf32), (void**)&r.vis);
        if (!rc && d.pcal)
            rc = fx_pcal_create_inputs(r.ctx, d.bandwidth_hz, d.pcal_spacing_hz, d.pcal_offset_hz, 0, d.ninp,
                                       &r.pcal);
    }
    if (!rc && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) rc = FX_ERR_HIP;
    for (int s = 0; s < nsub && !rc; s++) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) rc = FX_ERR_HIP;
        b->join.push_back(e);
    }
    if (!rc && b->use_graph) {
        // capture: fork the caller's stream to every subband stream and join back
        hipStream_t s0 = fx_ctx_hip_stream(ctx);
        if (hipStreamSynchronize(s0) != hipSuccess ||
            hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = FX_ERR_HIP;
        } else {
            int crc = hipEventRecord(b->fork, s0) == hipSuccess ? FX_OK : FX_ERR_HIP;
            for (int s = 0; s < nsub && !crc; s++) {
                Subband& r = b->sb[(size_t)s];
                hipStream_t ss = fx_ctx_hip_stream(r.ctx);
                if (hipStreamWaitEvent(ss, b->fork, 0) != hipSuccess) crc = FX_ERR_HIP;
                if (!crc) crc = enqueue(r);
                if (!crc && hipEventRecord(b->join[(size_t)s], ss) != hipSuccess) crc = FX_ERR_HIP;
                if (!crc && hipStreamWaitEvent(s0, b->join[(size_t)s], 0) != hipSuccess) crc = FX_ERR_HIP;
            }
            hipGraph_t g = nullptr;
            const bool ended = hipStreamEndCapture(s0, &g) == hipSuccess;
            if (crc || !ended) {
                if (g) hipGraphDestroy(g);
                rc = crc ? crc : FX_ERR_HIP;
                if (!crc) fx_set_error("fx_batch_create: graph capture failed");
            } else {
                b->graph = g;
                if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
                    fx_set_error("fx_batch_create: graph instantiation failed");
                    rc = FX_ERR_HIP;
                }
            }
        }
        hipGetLastError();
    }
    if (rc) {
        fx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return FX_OK;
}

extern "C" int fx_batch_run(fx_batch* b)
{
    FX_REQUIRE(b, FX_ERR_ARG, "null batch");
    hipStream_t s0 = fx_ctx_hip_stream(b->ctx);
    if (b->use_graph) {
        FX_HIP_TRY(hipGraphLaunch(b->exec, s0));
        return FX_OK;
    }
    // one launch at a time: every subband after the previous one's work
    FX_HIP_TRY(hipEventRecord(b->fork, s0));
    for (size_t s = 0; s < b->sb.size(); s++) {
        Subband& r = b->sb[s];
        hipStream_t ss = fx_ctx_hip_stream(r.ctx);
        FX_HIP_TRY(hipStreamWaitEvent(ss, s ? b->join[s - 1] : b->fork, 0));
        int rc = enqueue(r);
        if (rc) return rc;
        FX_HIP_TRY(hipEventRecord(b->join[s], ss));
    }
    FX_HIP_TRY(hipStreamWaitEvent(s0, b->join.back(), 0));
    return FX_OK;
}

extern "C" int fx_batch_sync(fx_batch* b)
{
    FX_REQUIRE(b, FX_ERR_ARG, "null batch");
    FX_HIP_TRY(hipStreamSynchronize(fx_ctx_hip_stream(b->ctx)));
    if (b->d_ctrs) {
        int err = 0;
        FX_HIP_TRY(hipMemcpy(&err, b->d_ctrs + FX_PCTR_ERR, sizeof(int), hipMemcpyDeviceToHost));
        FX_REQUIRE(err == 0, FX_ERR_HIP, "fx_batch_sync: a work item's wait timed out; its results are incomplete");
    }
    return FX_OK;
}

extern "C" int fx_batch_persist_tuning(fx_batch* b, int s, int* xtile, int* chan_per_item, int* groups_per_item)
{
    FX_REQUIRE(b && s >= 0 && s < (int)b->sb.size(), FX_ERR_ARG, "fx_batch_persist_tuning: subband %d", s);
    FX_REQUIRE(b->mode >= FX_BATCH_PERSISTENT, FX_ERR_STATE, "fx_batch_persist_tuning: not a persistent batch");
    const Subband& r = b->sb[(size_t)s];
    if (xtile) *xtile = r.xtile;
    if (chan_per_item) *chan_per_item = r.cb;
    if (groups_per_item) *groups_per_item = r.gpi;
    return FX_OK;
}

extern "C" int fx_batch_persist_info(fx_batch* b, int* nitems, int* nblocks, size_t* lds_bytes)
{
    FX_REQUIRE(b && b->mode >= FX_BATCH_PERSISTENT, FX_ERR_ARG, "fx_batch_persist_info: not a persistent batch");
    if (nitems) *nitems = b->nitems;
    if (nblocks) *nblocks = b->nblocks;
    if (lds_bytes) *lds_bytes = b->lds;
    return FX_OK;
}

extern "C" const fx_cf32* fx_batch_vis(fx_batch* b, int s)
{
    if (!b || s < 0 || s >= (int)b->sb.size()) return nullptr;
    return b->sb[(size_t)s].vis;
}

extern "C" int fx_batch_pcal(fx_batch* b, int s, int i, fx_cf64* tones, int* ntones)
{
    FX_REQUIRE(b && s >= 0 && s < (int)b->sb.size(), FX_ERR_ARG, "fx_batch_pcal: subband %d", s);
    Subband& r = b->sb[(size_t)s];
    FX_REQUIRE(r.pcal && i >= 0 && i < r.d.ninp, FX_ERR_ARG, "fx_batch_pcal: input %d has no phase-cal", i);
    int nt = 0;
    fx_pcal_info(r.pcal, nullptr, nullptr, &nt);
    if (ntones) *ntones = nt;
    if (!tones) return FX_OK;
    return fx_pcal_tones_input(r.pcal, i, tones, nullptr);
}

extern "C" int fx_batch_tuning(fx_batch* b, int s, long* sub, int* xtile, long* chunk)
{
    FX_REQUIRE(b && s >= 0 && s < (int)b->sb.size(), FX_ERR_ARG, "fx_batch_tuning: subband %d", s);
    const Subband& r = b->sb[(size_t)s];
    if (sub) *sub = r.sub;
    if (xtile) *xtile = r.xtile;
    if (chunk) *chunk = r.chunk;
    return FX_OK;
}

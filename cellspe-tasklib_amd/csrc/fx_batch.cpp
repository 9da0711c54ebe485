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
};

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
        if (s.vis) fx_free(s.ctx, s.vis);
        if (s.ctx) fx_ctx_destroy(s.ctx);
    }
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
    fx_batch* b = new fx_batch;
    b->ctx = ctx;
    b->prec = prec;
    b->use_graph = use_graph != 0;
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

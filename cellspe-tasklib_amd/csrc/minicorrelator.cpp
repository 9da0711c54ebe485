// minicorrelator.cpp -- minicorr_run(): the minicorrelator control-block
// drop-in (include/minicorrelator_fx.h) on the device F-X engine.
//
// The SPU program's per-period loop (minicorrelator_spu.c:326-481: unpack,
// rotate, FFT, exchange, three-baseline MAC, dump) becomes: stage every
// station's raw stream in HBM, then per integration block run the correlator
// (F-stage + X-engine over ffts_to_integrate spectra) and write the dump into
// the caller's buffers in the reference's split-plane layout
// (check_baseline_integration_period, minicorrelator_spu.c:835-866).
#include "fx_internal.h"

#include "minicorrelator_fx.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Run {
    fx_ctx* ctx = nullptr;
    fx_corr* corr = nullptr;
    void* words = nullptr;
    void* vis = nullptr;
    void* mult = nullptr;
    void* ev0 = nullptr;
    void* ev1 = nullptr;
    ~Run()
    {
        if (corr) fx_corr_destroy(corr);
        if (ctx) {
            if (words) fx_free(ctx, words);
            if (vis) fx_free(ctx, vis);
            if (mult) fx_free(ctx, mult);
            if (ev0) fx_event_destroy(ev0);
            if (ev1) fx_event_destroy(ev1);
            fx_ctx_destroy(ctx);
        }
    }
};

int baseline_index(int i, int j, int nst)
{
    // canonical i < j order: (0,1), (0,2), ..., (0,nst-1), (1,2), ...
    return i * nst - i * (i + 1) / 2 + (j - i - 1);
}

} // namespace

extern "C" int minicorr_run(const control_block* cb_in, int nst, minicorr_stats* stats)
{
    FX_REQUIRE(cb_in != nullptr, -1, "minicorr_run: null control blocks");
    FX_REQUIRE(nst >= 2 && nst <= NUM_SPE_THREADS, -1, "minicorr_run: %d stations (2..%d)", nst,
               NUM_SPE_THREADS);
    // the type is 128-byte aligned (the reference's SPU DMA needs it), but a
    // host caller -- an FFI binding -- need not align the array: read the
    // blocks as bytes into an aligned copy
    std::vector<control_block> cbv((size_t)nst);
    memcpy(cbv.data(), reinterpret_cast<const void*>(reinterpret_cast<uintptr_t>(cb_in)),
           sizeof(control_block) * (size_t)nst);
    const control_block* cb = cbv.data();
    const unsigned T = cb[0].ffts_total;
    const unsigned K = cb[0].ffts_to_integrate;
    FX_REQUIRE(K >= 1 && T >= K, -1, "minicorr_run: ffts_total=%u ffts_to_integrate=%u", T, K);
    for (int s = 0; s < nst; s++) {
        FX_REQUIRE(cb[s].spe_num == s, -1, "minicorr_run: control block %d has spe_num %d", s, cb[s].spe_num);
        FX_REQUIRE(cb[s].ffts_total == T && cb[s].ffts_to_integrate == K, -1,
                   "minicorr_run: station %d disagrees on ffts_total/ffts_to_integrate", s);
        FX_REQUIRE(cb[s].rawdata_src.p && cb[s].autocorrelation_out.p, -1,
                   "minicorr_run: station %d buffers missing", s);
    }
    const int nbl = nst * (nst - 1) / 2;
    for (int b = 0; b < nbl; b++)
        FX_REQUIRE(cb[0].baselines_out[b].p != nullptr, -1, "minicorr_run: baseline %d buffer missing", b);

    const int C = N_FFT / 2;
    const size_t nwords = (size_t)T * N_FFT / 16;
    const long P = (long)nst * (nst + 1) / 2;
    Run r;
    int rc = fx_ctx_create(0, &r.ctx);
    if (rc) return -1;
    if (fx_malloc(r.ctx, (size_t)nst * nwords * 4, &r.words) ||
        fx_malloc(r.ctx, (size_t)P * C * sizeof(fx_cf32), &r.vis))
        return -1;
    for (int s = 0; s < nst; s++)
        if (fx_memcpy_h2d(r.ctx, (char*)r.words + (size_t)s * nwords * 4, cb[s].rawdata_src.p, nwords * 4))
            return -1;
    const long chunk = std::min<long>(4096, ((long)K + 7) / 8 * 8);
    fx_corr_config cfg = {nst, N_FFT, chunk, FX_PREC_F32, 0};
    if (fx_corr_create(r.ctx, &cfg, &r.corr)) return -1;
    if (cb[0].fmultipliers_src.p) {
        // common per-channel phase (turns) applied to every station
        const float* f = (const float*)cb[0].fmultipliers_src.p;
        std::vector<fx_cf32> m((size_t)nst * C);
        for (int s = 0; s < nst; s++)
            for (int k = 0; k < C; k++) {
                double cyc = (double)f[k];
                cyc -= std::rint(cyc);
                m[(size_t)s * C + k] = {(float)std::cos(2.0 * M_PI * cyc), (float)std::sin(2.0 * M_PI * cyc)};
            }
        if (fx_malloc(r.ctx, m.size() * sizeof(fx_cf32), &r.mult) ||
            fx_memcpy_h2d(r.ctx, r.mult, m.data(), m.size() * sizeof(fx_cf32)) ||
            fx_corr_set_multipliers(r.corr, (const fx_cf32*)r.mult))
            return -1;
    }
    if (fx_event_create(r.ctx, &r.ev0) || fx_event_create(r.ctx, &r.ev1)) return -1;

    std::vector<fx_cf32> host((size_t)P * C);
    const unsigned ndump = T / K;   // a dump is written only for complete blocks
    float dev_ms = 0.f;
    for (unsigned b = 0; b < ndump; b++) {
        fx_event_record(r.ctx, r.ev0);
        if (fx_corr_reset(r.corr) || fx_corr_process(r.corr, (const uint32_t*)r.words, nwords, (long)b * K, K) ||
            fx_corr_dump(r.corr, (fx_cf32*)r.vis))
            return -1;
        fx_event_record(r.ctx, r.ev1);
        if (fx_memcpy_d2h(r.ctx, host.data(), r.vis, host.size() * sizeof(fx_cf32))) return -1;
        float ms = 0.f;
        fx_event_elapsed_ms(r.ev0, r.ev1, &ms);
        dev_ms += ms;
        for (int i = 0, p = 0; i < nst; i++)
            for (int j = i; j < nst; j++, p++) {
                float* dst = (float*)(i == j ? cb[i].autocorrelation_out.p
                                             : cb[0].baselines_out[baseline_index(i, j, nst)].p);
                dst += (size_t)b * 2 * N_FFT;
                float* re = dst;
                float* im = dst + N_FFT;
                const fx_cf32* v = &host[(size_t)p * C];
                for (int k = 0; k < C; k++) {
                    re[k] = v[k].re;
                    im[k] = v[k].im;
                }
                re[C] = 0.f;
                im[C] = 0.f;
                for (int k = C + 1; k < N_FFT; k++) {
                    re[k] = v[N_FFT - k].re;
                    im[k] = -v[N_FFT - k].im;
                }
            }
    }
    if (stats) {
        stats->device_ms = dev_ms;
        stats->spectra = (unsigned long long)ndump * K;
        stats->dumps = ndump;
    }
    return 0;
}

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
    // persistent modes: the whole integration's channelized samples (and
    // natural rows, slice sums), the device phase models, and the tuning:
    // natural rows + corner turns, channels per X item, F groups per item,
    // time slices of the X items
    uint32_t* chan = nullptr;
    uint32_t* nat = nullptr;
    float2* part = nullptr;
    size_t part_bytes = 0;
    fx_phase_model* pm_dev = nullptr;
    int ipad = 0;
    long ntp = 0;
    int pnat = 0, cb = 2, gpi = 4, nts = 1, xt = 64;
    // tuned launch times of the subband's F, CT and X items (ms, as one
    // subband among many): the fused table's interleaving weights
    float f_ms = 1.f, ct_ms = 0.f, x_ms = 1.f;
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
    int phase_end[4] = {0, 0, 0, 0};   // phased: launches over items [0,e0), [e0,e1), [e1,e2), [e2,e3)
    int nblocks = 0;
    size_t lds = 0;
};

namespace {

// ---------------------------------------------------------------------------
// Persistent modes
// ---------------------------------------------------------------------------
constexpr long kPartMax = 2L << 30;   // bytes of slice sums a subband may hold

long round_up(long n, long m) { return (n + m - 1) / m * m; }

// spectra per time slice (a multiple of the X stage) and the slices that gives
long slice_len(long ntp, int nts) { return round_up((ntp + nts - 1) / nts, FX_PXKT); }
int slices_of(long ntp, int nts) { return (int)((ntp + slice_len(ntp, nts) - 1) / slice_len(ntp, nts)); }

size_t part_bytes_of(const fx_subband_desc& d, int nts)
{
    const size_t P = (size_t)d.ninp * (d.ninp + 1) / 2, C = (size_t)d.nfft / 2;
    return nts > 1 ? (size_t)nts * P * C * sizeof(float2) : 0;
}

// the device view of subband r with its current tuning
void fill_psub(const Subband& r, fx_psub& S)
{
    const fx_subband_desc& d = r.d;
    S = fx_psub{};
    S.words = d.d_words;
    S.stride = d.stride_words;
    S.nfft = d.nfft;
    S.ninp = d.ninp;
    S.nt = d.nspectra;
    S.pm = r.pm_dev;
    // a 2-bit stream through the level table has E[x^2] ~ 4.316: the FFT
    // output's sigma sqrt(N * 4.316) maps to 1 before the FP16 cast
    S.sc = (float)(1.0 / sqrt((double)d.nfft * 4.316));
    S.ipad = r.ipad;
    S.chan = r.chan;
    S.nat = r.pnat ? r.nat : nullptr;
    S.vis = (float2*)r.vis;
    S.nts = r.nts;
    S.part = r.nts > 1 ? r.part : nullptr;
    S.ntp = r.ntp;
    S.tslice = slice_len(r.ntp, r.nts);
    if (r.pcal) fx_pcal_device_state(r.pcal, &S, nullptr);
}

int upload_psub(fx_batch* b, int s)
{
    fx_psub S;
    fill_psub(b->sb[(size_t)s], S);
    return fx_memcpy_h2d(b->ctx, b->d_subs + s, &S, sizeof(S));
}

// Work table of the persistent modes.  Counters FX_PCTR_FIRST +
// FX_PCTR_PER_SUB s + {0, 1, 2, 3} count the finished F, CT, X and PCAL items
// of subband s.  fused (FX_BATCH_PERSISTENT): the folds, then the subbands'
// F, CT and X items staggered (subband k's F items, k-1's corner turns,
// k-2's X items, ...) so that an item's inputs are mostly done when it is
// taken, then the slice reductions and the tones; phased: PCAL + F, CT, X,
// RED + TONE as four ranges.
int build_table(fx_batch* b, bool fused, std::vector<fx_pitem>& items, int* phase_end)
{
    const int ns = (int)b->sb.size();
    std::vector<int> nF((size_t)ns, 0), nC((size_t)ns, 0), nX((size_t)ns, 0), nP((size_t)ns, 0);
    auto ctr = [](int s, int k) { return FX_PCTR_FIRST + FX_PCTR_PER_SUB * s + k; };
    auto item = [](int type, int sb, long a, long bb, int c, int d, int wc, int wv, int dc) {
        fx_pitem it;
        it.type = type;
        it.sb = sb;
        it.a = a;
        it.b = bb;
        it.c = c;
        it.d = d;
        it.wait_ctr = wc;
        it.wait_val = wv;
        it.done_ctr = dc;
        it.xt = 64;
        return it;
    };
    auto pcal_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (!r.pcal) return;
        fx_psub S{};
        fx_pcal_device_state(r.pcal, &S, nullptr);
        const long nbin = S.nfold / 16;
        const long wpi = nbin * std::max<long>(1, 262144 / nbin);   // about 4M samples per fold item
        const long nw = r.d.nspectra * r.d.nfft / 16;
        for (int y = 0; y < r.d.ninp; y++)
            for (long w0 = 0; w0 < nw; w0 += wpi) {
                items.push_back(item(FX_PI_PCAL, s, y, w0, (int)wpi, 0, -1, 0, ctr(s, 3)));
                nP[(size_t)s]++;
            }
    };
    auto f_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        const long F = fx_persist_group_items(r.d.nfft);
        const long G = (r.d.nspectra * r.d.ninp + F - 1) / F;
        // fused: twice the tuned groups per item.  Each finished item's
        // agent-scope release writes back its XCD's L2 (the XCDs' L2s are not
        // coherent), a cost the tuner's phased timings do not see; halving
        // the F items' count and releases took the 32-subband batch from
        // 22.2-22.9 to 20.8 ms (4x: 20.8-21.4 ms, 8x: 23.5 ms, imbalance)
        const long gpi = fused ? 2 * r.gpi : r.gpi;
        for (long g0 = 0; g0 < G; g0 += gpi) {
            items.push_back(item(FX_PI_F, s, g0, std::min<long>(gpi, G - g0), 0, 0, -1, 0, ctr(s, 0)));
            nF[(size_t)s]++;
        }
    };
    auto ct_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (!r.pnat) return;
        const int nblk = (r.d.nfft / 2 / 64) * ((r.d.ninp + 63) / 64);
        for (long t0 = 0; t0 < r.d.nspectra; t0 += FX_PCT_T)
            for (int k = 0; k < nblk; k++) {
                items.push_back(item(FX_PI_CT, s, t0, FX_PCT_T, k, 0, ctr(s, 0), nF[(size_t)s], ctr(s, 1)));
                nC[(size_t)s]++;
            }
    };
    auto x_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        const int nb = r.xt == 128 ? 1 : (r.d.ninp + 63) / 64;
        const long C = r.d.nfft / 2;
        const int wc = r.pnat ? ctr(s, 1) : ctr(s, 0);
        const int wv = r.pnat ? nC[(size_t)s] : nF[(size_t)s];
        // X items are counted (a release each) only for the slice reductions
        const int dc = r.nts > 1 ? ctr(s, 2) : -1;
        for (int sl = 0; sl < r.nts; sl++)
            for (int t = 0; t < nb * (nb + 1) / 2; t++)
                for (long c0 = 0; c0 < C; c0 += r.cb) {
                    fx_pitem it = item(FX_PI_X, s, c0, std::min<long>(r.cb, C - c0), t, sl, wc, wv, dc);
                    it.xt = r.xt;
                    items.push_back(it);
                    nX[(size_t)s]++;
                }
    };
    auto red_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (r.nts < 2) return;
        const long P = (long)r.d.ninp * (r.d.ninp + 1) / 2, C = r.d.nfft / 2;
        const long pch = std::max<long>(1, 262144 / C);   // about 256K visibilities per item
        for (long p0 = 0; p0 < P; p0 += pch)
            items.push_back(item(FX_PI_RED, s, p0, std::min<long>(pch, P - p0), 0, 0, ctr(s, 2), nX[(size_t)s], -1));
    };
    auto t_items = [&](int s) {
        const Subband& r = b->sb[(size_t)s];
        if (!r.pcal) return;
        for (int y = 0; y < r.d.ninp; y++)
            items.push_back(item(FX_PI_TONE, s, y, 0, 0, 0, ctr(s, 3), nP[(size_t)s], -1));
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
    if (fused) {
        // round k: the folds and F items of subband k, the corner turns of
        // k - 2, the X items of k - 2 (direct stores) or k - 4 (after corner
        // turns): an item's inputs were made a whole round earlier, so it
        // rarely waits.  The streams of a round are merged so that each is the
        // same fraction done at every point of the round (by the tuned times),
        // so the workgroups run VALU-bound F work beside HBM-bound X work.
        auto xround = [&](int k) { return k + (b->sb[(size_t)order[(size_t)k]].pnat ? 4 : 2); };
        for (int k = 0; k < ns + 4; k++) {
            std::vector<fx_pitem> st[5];
            double cost[5] = {0, 0, 0, 0, 0};
            const size_t base = items.size();
            auto take = [&](int q, double c) {
                st[q].assign(items.begin() + (long)base, items.end());
                items.resize(base);
                cost[q] = c;
            };
            if (k < ns) {
                const int s = order[(size_t)k];
                pcal_items(s);
                take(0, 0.1 * b->sb[(size_t)s].f_ms);
                f_items(s);
                take(1, b->sb[(size_t)s].f_ms);
            }
            if (k >= 2 && k - 2 < ns) {
                const int s = order[(size_t)k - 2];
                ct_items(s);
                take(2, b->sb[(size_t)s].ct_ms);
            }
            for (int j = std::max(0, k - 4); j <= k - 2 && j < ns; j++) {
                if (xround(j) != k) continue;
                const int s = order[(size_t)j];
                x_items(s);
                take(j == k - 2 ? 3 : 4, b->sb[(size_t)s].x_ms);
            }
            size_t pos[5] = {0, 0, 0, 0, 0};
            for (;;) {
                int pick = -1;
                double best = 2.0;
                for (int q = 0; q < 5; q++) {
                    if (pos[q] >= st[q].size()) continue;
                    // fraction of the stream already placed (streams without
                    // a time estimate go first)
                    const double frac = cost[q] > 0 ? (double)pos[q] / (double)st[q].size() : -1.0;
                    if (frac < best) {
                        best = frac;
                        pick = q;
                    }
                }
                if (pick < 0) break;
                items.push_back(st[pick][pos[pick]++]);
            }
        }
        phase_end[0] = phase_end[1] = phase_end[2] = (int)items.size();
    } else {
        for (int s : order) pcal_items(s);
        for (int s : order) f_items(s);
        phase_end[0] = (int)items.size();
        for (int s : order) ct_items(s);
        phase_end[1] = (int)items.size();
        for (int s : order) x_items(s);
        phase_end[2] = (int)items.size();
    }
    for (int s : order) red_items(s);
    for (int s : order) t_items(s);
    phase_end[3] = (int)items.size();
    // every wait names items that come earlier in the table: the workgroups
    // that took them are running when the wait starts, so it ends
    std::vector<int> seen((size_t)(FX_PCTR_FIRST + FX_PCTR_PER_SUB * ns), 0);
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
// one fused launch or four phased ones
int enqueue_persist(fx_batch* b, hipStream_t s, const fx_pitem* d_items, const int* phase_end, bool fused)
{
    if (fxk_fill32(s, b->d_ctrs, 0u, (size_t)b->nctrs) != FX_OK) return FX_ERR_HIP;
    for (const Subband& r : b->sb) {
        if (!r.pcal) continue;
        fx_psub S{};
        fx_pcal_device_state(r.pcal, &S, nullptr);
        if (fxk_fill32(s, S.pacc, 0u, (size_t)S.nfold * r.d.ninp * 2) != FX_OK) return FX_ERR_HIP;
    }
    const float2* tw = fx_ctx_twiddles(b->ctx);
    const int nb = b->nblocks;
    if (fused)
        return fxk_batch_persistent(s, tw, b->d_subs, d_items, 0, phase_end[3], b->d_ctrs, 0, nb, 1, b->lds);
    int i0 = 0;
    for (int ph = 0; ph < 4; ph++) {
        int rc = fxk_batch_persistent(s, tw, b->d_subs, d_items, i0, phase_end[ph], b->d_ctrs, ph, nb, 0, b->lds);
        if (rc) return rc;
        i0 = phase_end[ph];
    }
    return FX_OK;
}

bool lookup_persist(const char* file, const fx_subband_desc& d, int copies, Subband& r)
{
    if (!file) return false;
    FILE* f = fopen(file, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        int n, i, bl, na, c, g, ts, xt;
        long t;
        float fms, cms, xms;
        if (sscanf(line, "P %d %d %ld %d %d %d %d %d %d %f %f %f", &n, &i, &t, &bl, &na, &c, &g, &ts, &xt, &fms, &cms,
                   &xms) == 12 &&
            n == d.nfft && i == d.ninp && t == d.nspectra && bl == copies) {
            r.pnat = na;
            r.cb = c;
            r.gpi = g;
            r.nts = ts;
            r.xt = xt;
            r.f_ms = fms;
            r.ct_ms = cms;
            r.x_ms = xms;
            found = true;
        }
    }
    fclose(f);
    return found;
}

// defaults: natural rows from 4096-point transforms, channel pairs of X
// items of about 2048 spectra of channels, 4 F groups per item, one slice
void persist_defaults(Subband& r)
{
    const fx_subband_desc& d = r.d;
    r.pnat = d.nfft >= 4096 ? 1 : 0;
    r.cb = (int)std::max<long>(2, std::min<long>(d.nfft / 2, 2048 / std::max<long>(1, d.nspectra)) / 2 * 2);
    r.gpi = 4;
    r.nts = 1;
    r.xt = d.ninp > 64 && d.ninp <= 128 ? 128 : 64;
    if (r.xt == 128) r.cb = 1;
}

// table of one subband only (the autotuner's timing runs), phased
int build_table_one(fx_batch* b, int s, std::vector<fx_pitem>& items, int* phase_end)
{
    fx_batch one;
    one.sb.push_back(b->sb[(size_t)s]);
    int rc = build_table(&one, false, items, phase_end);
    for (fx_pitem& it : items) {
        it.sb = s;
        if (it.done_ctr >= 0) it.done_ctr += FX_PCTR_PER_SUB * s;
        if (it.wait_ctr >= 0) it.wait_ctr += FX_PCTR_PER_SUB * s;
    }
    one.sb.clear();   // owned by b
    return rc;
}

// Time launches [ph0, ph1] of subband s's phased table with its current
// tuning, in ms per copy (the second of two runs): every launch holds `rep`
// copies of the subband's items on all the batch's workgroups, so the
// subband runs on about its share of a loaded device (the copies write the
// same values; the real runs clear and rewrite everything).
int time_phases(fx_batch* b, int s, int ph0, int ph1, int rep, float* ms)
{
    std::vector<fx_pitem> one, items;
    int pe1[4], pe[4];
    int rc = build_table_one(b, s, one, pe1);
    if (!rc) rc = upload_psub(b, s);
    if (rc) return rc;
    for (int ph = 0, i0 = 0; ph < 4; ph++) {
        for (int r = 0; r < rep; r++) items.insert(items.end(), one.begin() + i0, one.begin() + pe1[ph]);
        pe[ph] = (int)items.size();
        i0 = pe1[ph];
    }
    fx_pitem* d = nullptr;
    rc = fx_malloc(b->ctx, items.size() * sizeof(fx_pitem), (void**)&d);
    if (rc) return rc;
    rc = fx_memcpy_h2d(b->ctx, d, items.data(), items.size() * sizeof(fx_pitem));
    void *e0 = nullptr, *e1 = nullptr;
    if (!rc) rc = fx_event_create(b->ctx, &e0);
    if (!rc) rc = fx_event_create(b->ctx, &e1);
    hipStream_t st = fx_ctx_hip_stream(b->ctx);
    const float2* tw = fx_ctx_twiddles(b->ctx);
    for (int run = 0; run < 2 && !rc; run++) {
        rc = fxk_fill32(st, b->d_ctrs, 0u, (size_t)b->nctrs);
        // the earlier launches once untimed (their outputs feed the timed ones)
        int i0 = 0;
        for (int ph = 0; ph < ph0 && !rc; ph++) {
            rc = fxk_batch_persistent(st, tw, b->d_subs, d, i0, pe[ph], b->d_ctrs, ph, b->nblocks, 0, b->lds);
            i0 = pe[ph];
        }
        if (!rc) rc = fx_event_record(b->ctx, e0);
        for (int ph = ph0; ph <= ph1 && !rc; ph++) {
            rc = fxk_batch_persistent(st, tw, b->d_subs, d, i0, pe[ph], b->d_ctrs, ph, b->nblocks, 0, b->lds);
            i0 = pe[ph];
        }
        if (!rc) rc = fx_event_record(b->ctx, e1);
        if (!rc) rc = fx_event_elapsed_ms(e0, e1, ms);
    }
    *ms /= (float)rep;
    if (e0) fx_event_destroy(e0);
    if (e1) fx_event_destroy(e1);
    fx_ctx_sync(b->ctx);
    fx_free(b->ctx, d);
    return rc;
}

// Autotuner of the persistent modes for subband s, timing that subband's
// items in `rep` concurrent copies (its share of the device in the batch):
// natural rows + corner turns or direct stores (transforms from 2048 points)
// and F groups per item from {4, 8, 16, 32} on the F (+ CT) launches, then
// the X tile width (one 128-input tile for 65 to 128 inputs, or 64-input
// tiles) and channels per X item from {1 or 2, ..., 64} on the X launch, then
// time slices from {2, 4, 8, 16} on the X + reduction launches.  The choice
// and the F, CT and X launch times are appended to tune_file as
// "P nfft ninp nspectra copies nat cb gpi nts xt f_ms ct_ms x_ms".
int autotune_persist(fx_batch* b, int s, int rep, const char* tune_file)
{
    Subband& r = b->sb[(size_t)s];
    const fx_subband_desc& d = r.d;
    float best = 1e30f, ms = 0.f;
    int rc = FX_OK;
    // F: natural rows + corner turns or direct stores, groups per item
    int best_n = r.pnat, best_g = r.gpi;
    for (int na = 0; na <= (r.nat ? 1 : 0); na++)
        for (int g : {4, 8, 16, 32}) {
            r.pnat = na;
            r.gpi = g;
            if ((rc = time_phases(b, s, 0, 1, rep, &ms))) return rc;
            if (ms < best) {
                best = ms;
                best_n = na;
                best_g = g;
            }
        }
    r.pnat = best_n;
    r.gpi = best_g;
    if ((rc = time_phases(b, s, 0, 0, rep, &r.f_ms))) return rc;
    r.ct_ms = 0.f;
    if (r.pnat && (rc = time_phases(b, s, 1, 1, rep, &r.ct_ms))) return rc;
    // X: one 128-input tile or 64-input tiles (65 to 128 inputs), channels
    // per item, then time slices
    best = 1e30f;
    r.nts = 1;
    int best_x = r.xt, best_c = r.cb;
    const long C = d.nfft / 2;
    for (int xt : {64, 128}) {
        if (xt == 128 && !(d.ninp > 64 && d.ninp <= 128)) continue;
        for (int cb = xt == 128 ? 1 : 2; cb <= 64 && cb <= C; cb *= 2) {
            r.xt = xt;
            r.cb = cb;
            if ((rc = time_phases(b, s, 2, 3, rep, &ms))) return rc;
            if (ms < best) {
                best = ms;
                best_x = xt;
                best_c = cb;
            }
        }
    }
    r.xt = best_x;
    r.cb = best_c;
    int best_t = 1, last = 1;
    for (int ts : {2, 4, 8, 16}) {
        const int eff = slices_of(r.ntp, ts);
        if (eff <= last || part_bytes_of(d, eff) > r.part_bytes) continue;
        last = eff;
        r.nts = eff;
        if ((rc = time_phases(b, s, 2, 3, rep, &ms))) return rc;
        if (ms < best) {
            best = ms;
            best_t = eff;
        }
    }
    r.nts = best_t;
    r.x_ms = best;
    if (tune_file) {
        FILE* f = fopen(tune_file, "a");
        if (f) {
            fprintf(f, "P %d %d %ld %d %d %d %d %d %d %.4f %.4f %.4f\n", d.nfft, d.ninp, d.nspectra, rep, r.pnat, r.cb,
                    r.gpi, r.nts, r.xt, r.f_ms, r.ct_ms, r.x_ms);
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
    size_t lds = std::max({fx_persist_lds(FX_PI_TONE, 0, false), fx_persist_lds(FX_PI_CT, 0, false),
                           fx_persist_lds(FX_PI_X, 0, false), fx_persist_lds(FX_PI_X, 1, false)});
    for (int s = 0; s < nsub && !rc; s++) {
        Subband& r = b->sb[(size_t)s];
        r.d = sb[s];
        const fx_subband_desc& d = r.d;
        const bool pow2 = d.nfft >= 256 && d.nfft <= 16384 && (d.nfft & (d.nfft - 1)) == 0;
        if (!pow2 || d.ninp < 1 || d.nspectra < 1 || !d.d_words || d.stride_words % 2 ||
            ((uintptr_t)d.d_words % 8) || d.stride_words < (size_t)d.nspectra * d.nfft / 16 ||
            (double)d.nspectra * d.ninp >= 2147483647.0 || d.nspectra >= (1L << 30)) {
            fx_set_error("fx_batch_create: subband %d: nfft %d, %d inputs, %ld spectra, word stride %zu", s, d.nfft,
                         d.ninp, d.nspectra, d.stride_words);
            rc = FX_ERR_ARG;
            break;
        }
        r.chunk = d.nspectra;
        persist_defaults(r);
        r.ipad = (int)round_up(d.ninp, 4);
        r.ntp = round_up(d.nspectra, FX_PXKT);
        const size_t C = (size_t)d.nfft / 2, P = (size_t)d.ninp * (d.ninp + 1) / 2;
        // chan: the pad inputs and spectra stay zero; 1 KiB of slack for the
        // X items' 64-input slab rows past the last channel's last row
        const size_t chan_bytes = C * (size_t)r.ntp * (size_t)r.ipad * 4 + 1024;
        rc = fx_malloc(ctx, chan_bytes, (void**)&r.chan);
        if (!rc) rc = fx_memset(ctx, r.chan, 0, chan_bytes);
        if (!rc) rc = fx_malloc(ctx, P * C * sizeof(fx_cf32), (void**)&r.vis);
        if (!rc && d.nfft >= 2048)
            rc = fx_malloc(ctx, C * (size_t)d.nspectra * (size_t)r.ipad * 4, (void**)&r.nat);
        if (!rc && d.pm) {
            rc = fx_malloc(ctx, sizeof(fx_phase_model) * (size_t)d.ninp, (void**)&r.pm_dev);
            if (!rc) rc = fx_memcpy_h2d(ctx, r.pm_dev, d.pm, sizeof(fx_phase_model) * (size_t)d.ninp);
        }
        if (!rc && d.pcal) {
            fx_psub S{};
            rc = fx_pcal_create_inputs(ctx, d.bandwidth_hz, d.pcal_spacing_hz, d.pcal_offset_hz, 0, d.ninp, &r.pcal);
            if (!rc) rc = fx_pcal_device_state(r.pcal, &S, nullptr);
            if (!rc && (S.nfold % 16 || S.nfold > 4096)) {
                fx_set_error("fx_batch_create: subband %d: a phase-cal fold of %d bins (the persistent modes fold "
                             "multiples of 16 up to 4096)", s, S.nfold);
                rc = FX_ERR_UNSUPPORTED;
            }
            lds = std::max(lds, fx_persist_lds(FX_PI_PCAL, S.nfold, false));
        }
        lds = std::max(lds, fx_persist_lds(FX_PI_F, d.nfft, d.pm != nullptr));
    }
    const int ns = nsub;
    b->nctrs = FX_PCTR_FIRST + FX_PCTR_PER_SUB * ns;
    if (!rc) rc = fx_malloc(ctx, sizeof(fx_psub) * (size_t)ns, (void**)&b->d_subs);
    if (!rc) rc = fx_malloc(ctx, sizeof(int) * (size_t)b->nctrs, (void**)&b->d_ctrs);
    if (!rc) {
        b->lds = lds;
        b->nblocks = fxk_batch_persistent_blocks(lds);
        if (b->nblocks < 1) rc = FX_ERR_HIP;   // fx_last_error() says why
    }
    // each subband's share of the device: its part of the batch's work, at
    // least an eighth; tuned as that many concurrent copies of it
    std::vector<int> tcopies((size_t)ns, 1);
    if (!rc) {
        double tot = 0;
        std::vector<double> w((size_t)ns);
        for (int s = 0; s < ns; s++) {
            const fx_subband_desc& d = b->sb[(size_t)s].d;
            w[(size_t)s] = (double)d.nspectra * d.ninp * (2.5 * d.nfft * std::log2((double)d.nfft) + 4.0 * d.nfft * (d.ninp + 1));
            tot += w[(size_t)s];
        }
        for (int s = 0; s < ns; s++) {
            const double share = std::max(0.125, std::min(1.0, w[(size_t)s] / std::max(tot, 1.0)));
            tcopies[(size_t)s] = std::max(1, (int)std::lround(1.0 / share));
        }
    }
    std::vector<bool> tuned((size_t)ns, false);
    for (int s = 0; s < ns && !rc; s++) {
        Subband& r = b->sb[(size_t)s];
        if (tune_file) tuned[(size_t)s] = lookup_persist(tune_file, r.d, tcopies[(size_t)s], r);
        if (!r.nat) r.pnat = 0;
    }
    // tune each shape once (the tune file already holds the others); the
    // slice sums get a temporary buffer of the largest candidate that fits
    for (int s = 0; s < ns && !rc; s++) {
        if (!tune_file || tuned[(size_t)s]) continue;
        Subband& r = b->sb[(size_t)s];
        size_t pb = 0;
        for (int ts : {16, 8, 4, 2})
            if (!pb && part_bytes_of(r.d, slices_of(r.ntp, ts)) <= (size_t)kPartMax)
                pb = part_bytes_of(r.d, slices_of(r.ntp, ts));
        if (pb) rc = fx_malloc(ctx, pb, (void**)&r.part);
        r.part_bytes = rc ? 0 : pb;
        if (!rc) rc = autotune_persist(b, s, tcopies[(size_t)s], tune_file);
        if (r.part) fx_free(ctx, r.part);
        r.part = nullptr;
        r.part_bytes = 0;
        for (int q = s + 1; q < ns; q++) {
            Subband& o = b->sb[(size_t)q];
            if (!tuned[(size_t)q] && o.d.nfft == r.d.nfft && o.d.ninp == r.d.ninp && o.d.nspectra == r.d.nspectra &&
                tcopies[(size_t)q] == tcopies[(size_t)s]) {
                o.pnat = r.pnat;
                o.cb = r.cb;
                o.gpi = r.gpi;
                o.nts = r.nts;
                o.xt = r.xt;
                o.f_ms = r.f_ms;
                o.ct_ms = r.ct_ms;
                o.x_ms = r.x_ms;
                tuned[(size_t)q] = true;
            }
        }
    }
    // the chosen layouts: slice sums where the X items are sliced, natural
    // rows only where they are used
    for (int s = 0; s < ns && !rc; s++) {
        Subband& r = b->sb[(size_t)s];
        r.nts = slices_of(r.ntp, std::max(1, r.nts));
        if (r.nts > 1) {
            r.part_bytes = part_bytes_of(r.d, r.nts);
            rc = fx_malloc(ctx, r.part_bytes, (void**)&r.part);
        }
        if (!rc && !r.pnat && r.nat) {
            fx_free(ctx, r.nat);
            r.nat = nullptr;
        }
        if (!rc) rc = upload_psub(b, s);
    }
    std::vector<fx_pitem> items;
    if (!rc) rc = build_table(b, mode == FX_BATCH_PERSISTENT, items, b->phase_end);
    if (!rc) {
        b->nitems = (int)items.size();
        rc = fx_malloc(ctx, sizeof(fx_pitem) * items.size(), (void**)&b->d_items);
    }
    if (!rc) rc = fx_memcpy_h2d(ctx, b->d_items, items.data(), sizeof(fx_pitem) * items.size());
    if (!rc) rc = fx_ctx_sync(ctx);
    if (!rc) {
        // one graph: the clears and the launch(es)
        hipStream_t s0 = fx_ctx_hip_stream(ctx);
        if (hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = FX_ERR_HIP;
        } else {
            int crc = enqueue_persist(b, s0, b->d_items, b->phase_end, mode == FX_BATCH_PERSISTENT);
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
        if (s.nat) fx_free(own, s.nat);
        if (s.part) fx_free(own, s.part);
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

extern "C" int fx_batch_persist_tuning(fx_batch* b, int s, int* natural_rows, int* chan_per_item,
                                       int* groups_per_item, int* time_slices, int* xtile)
{
    FX_REQUIRE(b && s >= 0 && s < (int)b->sb.size(), FX_ERR_ARG, "fx_batch_persist_tuning: subband %d", s);
    FX_REQUIRE(b->mode >= FX_BATCH_PERSISTENT, FX_ERR_STATE, "fx_batch_persist_tuning: not a persistent batch");
    const Subband& r = b->sb[(size_t)s];
    if (natural_rows) *natural_rows = r.pnat;
    if (chan_per_item) *chan_per_item = r.cb;
    if (groups_per_item) *groups_per_item = r.gpi;
    if (time_slices) *time_slices = r.nts;
    if (xtile) *xtile = r.xt;
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

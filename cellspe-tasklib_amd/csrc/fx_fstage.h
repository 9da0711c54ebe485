// fx_fstage.h -- building blocks of the F kernels (2-bit unpack -> real FFT
// -> FP64-model rotation -> channelized store) shared by the F-stage kernels
// of fx_fft.hip and the persistent batch kernel of fx_persist.hip: the
// channelized sample formats, the per-item FP64 phasor tables, the LDS
// tables and the front of the transform (unpack through the last FFT pass).
#pragma once

#include "fx_internal.h"

#include <hip/hip_fp8.h>

#include <cstddef>
#include <type_traits>

namespace {

// ---------------------------------------------------------------------------
// F-stage.  Item g = t*I + i of a chunk (g = i*nt + t for FP8 output) is
// spectrum t_first + t of input i.  The kernel writes its natural output rows
// nat[g][c] (each transform's C channels contiguous, every store instruction
// a 256-byte run), and k_cornerturn re-tiles them through LDS into the
// X-engine layout out[c][g] ([c][t][i], or [c][i][t] for FP8).  Writing [c][t][i] straight from the transform
// scatters 4-16 byte fragments that the L2 writes back as partial lines.
// ---------------------------------------------------------------------------
template <typename T> struct ChanT;
template <> struct ChanT<float2> { typedef float2 V; };
template <> struct ChanT<__half2> { typedef uint32_t V; };
template <> struct ChanT<uint16_t> { typedef uint16_t V; };

template <typename T>
__device__ __forceinline__ typename ChanT<T>::V to_chan(float2 v);
template <>
__device__ __forceinline__ float2 to_chan<float2>(float2 v) { return v; }
template <>
__device__ __forceinline__ uint32_t to_chan<__half2>(float2 v)
{
    // the FP32 value first (an empty asm the compiler cannot see through), then
    // one rounding to FP16: no fused multiply-add-to-FP16 in some epilogues
    // and not in others
    asm volatile("" : "+v"(v.x), "+v"(v.y));
    __half2 h = __floats2half2_rn(v.x, v.y);
    return *reinterpret_cast<uint32_t*>(&h);
}
template <>
__device__ __forceinline__ uint16_t to_chan<uint16_t>(float2 v)
{
    __hip_fp8_e4m3 r(v.x), i(v.y);
    return (uint16_t)((uint16_t)r.__x | ((uint16_t)i.__x << 8));
}

// Phasor tables of the F kernels: channel k = PLO hi + lo of an item rotates
// by ph_hi[hi] * ph_lo[lo] = e^{2 pi i (k tau/N + fringe/2pi)}, both factors
// evaluated from the FP64 phase model (sincospi in FP64, one rounding to
// FP32), one complex multiply per channel.  The item's amplitude factor
// amp = 1/2 (of the real split) x the uniform channel scale x the input scale
// rides in ph_lo (or multiplies each channel when there is no phase model).
// The split keeps PLO + PHI (the FP64 sincospi evaluations per item) near
// its minimum 2 sqrt(M): PLO = 64 from M = 2048 up, 32 for M = 512 and 1024,
// 16 below (M = 128: 24 entries per item rather than 66).
template <int M>
struct PhaseTab {
    static constexpr int PLO = M >= 2048 ? 64 : M >= 512 ? 32 : M >= 16 ? 16 : M;
    static constexpr int PHI = M / PLO;
};

// The item bookkeeping comes first: a kernel without a phase model may give
// an FItems only the LDS up to ph_lo (fitems_head_bytes).
template <int M, int F>
struct FItems {
    int t[F], i[F];   // (spectrum, input) of item f; t = -1 past the end
    float amp[F];
    float2 ph_lo[F][PhaseTab<M>::PLO];
    float2 ph_hi[F][PhaseTab<M>::PHI];
};
template <int M, int F>
__host__ __device__ constexpr size_t fitems_head_bytes()
{
    using T = FItems<M, F>;
    return offsetof(T, ph_lo);
}

// Tables of the F kernels in LDS: the 16-entry unpack table, the second-pass
// twiddles tw2[r*16 + k] = W_256^{r k}, for M = 4096 the third-pass tables
// t3 = {W_4096^k, W_4096^{4k}, W_4096^{8k}} (k < 256), and the two-level
// table twl only where another pass needs it (the LDS of two workgroups per
// CU leaves room for t3 only without twl).
template <int M>
struct FTabs {
    static constexpr bool T3 = M == 4096;
    static constexpr bool TWL = !(M == 256 || M == 4096);
    float2 lut[16];
    float2 tw2[M >= 256 ? 256 : 1];
    float2 t3[T3 ? 768 : 1];
    float2 twl[TWL ? FX_TWL : 1];
};

// Fill the tables (visible after the first barrier of fstage_group).
template <int M, int NT>
__device__ __forceinline__ void fstage_consts(const float2* __restrict__ twg, FTabs<M>& tb, int tid)
{
    if (tid < 16) tb.lut[tid] = make_float2(lut2((uint32_t)tid & 3u), lut2((uint32_t)tid >> 2));
    if constexpr (M >= 256) {
        for (int e = tid; e < 256; e += NT) tb.tw2[e] = twg[(e >> 4) * (e & 15) * (FX_TW_N / 256)];
    }
    if constexpr (FTabs<M>::T3) {
        // W_4096^{m k} = T[m k FX_TW_N / 4096], m = 1, 4, 8
        for (int e = tid; e < 768; e += NT) {
            const int blk = e >> 8, k = e & 255;
            tb.t3[e] = twg[(k << (blk == 0 ? 3 : blk == 1 ? 5 : 6)) & (FX_TW_N - 1)];
        }
    }
    if constexpr (FTabs<M>::TWL) twl_load(tb.twl, twg, tid, NT);
}

// The same tables as register-held entries: load() issues the global reads
// of thread tid's entries, put() writes them (and the unpack table) to LDS,
// so a kernel's first group can have these reads in flight together with its
// packed words and phase models (one memory round trip before the first
// barrier instead of three).
template <int M, int NT>
struct FConstsPre {
    static constexpr int N2 = M >= 256 ? 256 : 0;
    static constexpr int N3 = FTabs<M>::T3 ? 768 : 0;
    static constexpr int NL = FTabs<M>::TWL ? FX_TWL : 0;
    static constexpr int TOT = N2 + N3 + NL;
    static constexpr int PER = (TOT + NT - 1) / NT;
    float2 v[PER > 0 ? PER : 1];
    // index into the master table of entry e (tw2, then t3, then twl)
    __device__ static __forceinline__ int src(int e)
    {
        if (e < N2) return (e >> 4) * (e & 15) * (FX_TW_N / 256);
        e -= N2;
        if (e < N3) {
            const int blk = e >> 8, k = e & 255;
            return (k << (blk == 0 ? 3 : blk == 1 ? 5 : 6)) & (FX_TW_N - 1);
        }
        e -= N3;
        return e < 256 ? e : (e - 256) << 8;
    }
    __device__ __forceinline__ void load(const float2* __restrict__ twg, int tid)
    {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = tid + p * NT;
            if (e < TOT) v[p] = twg[src(e)];
        }
    }
    __device__ __forceinline__ void put(FTabs<M>& tb, int tid) const
    {
        if (tid < 16) tb.lut[tid] = make_float2(lut2((uint32_t)tid & 3u), lut2((uint32_t)tid >> 2));
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = tid + p * NT;
            if (e < N2) {
                tb.tw2[e] = v[p];
            } else if (e < N2 + N3) {
                tb.t3[e - N2] = v[p];
            } else if (e < TOT) {
                tb.twl[e - N2 - N3] = v[p];
            }
        }
    }
};

// A value the caller knows to be the same in every lane, moved to scalar
// registers (the integer arithmetic on it then runs on the scalar unit).
__device__ __forceinline__ long uniform_long(long v)
{
    const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long)(((unsigned long)(unsigned)hi << 32) | (unsigned long)(unsigned)lo);
}

// (spectrum t, input i) of item g of a chunk of nt spectra x ninp inputs
// (g = t*I + i, or g = i*nt + t when IM); false past the chunk's last item.
template <bool IM>
__device__ __forceinline__ bool fitem(long g, long nt, int ninp, long nitems, long& t, int& i)
{
    t = IM ? g % nt : g / ninp;
    i = (int)(IM ? g / nt : g - t * ninp);
    return g < nitems;
}

// Front of the F kernels for the F items g0 .. g0+F-1: item bookkeeping, the
// FP64 phasor tables, the LDS FFT.  Item order: g = t*I + i, or g = i*nt + t
// when IM.  The packed words of the items are staged at the head of sh and
// the first radix-16 pass reads its 16 points per butterfly straight from
// them: complex point e of an item is the 4-bit field e%8 of word e/8, one
// lookup in the 16-entry table lut (both samples of the point).
// hook(std::integral_constant<int, p>) runs after FFT pass p
// (p < fft_npasses<M, 1>()), between the passes' barriers.  pre() runs first
// and issues the group's global reads into registers; stage(raw) then writes
// the group's packed words to raw (thread tid the WTH words from WTH*tid on)
// after the phasor tables are built, so the reads of words, phase models and
// tables are in flight together.  Every thread derives the (spectrum, input)
// of the items it needs itself: no barrier before stage.  The overload
// without pre/stage loads the words from `words`.
template <int M, int F, int NT, bool IM, bool ONEROW = false, typename Hook, typename Pre, typename Stage>
__device__ __forceinline__ void fstage_group(const uint32_t* __restrict__ words, size_t in_stride, int ninp,
                                             long t_first, long nt, const fx_phase_model* __restrict__ pm,
                                             float sc_uni, const float* __restrict__ sc_inp, long g0, float2* sh,
                                             const FTabs<M>& tb, FItems<M, F>& it, int tid, Hook&& hook,
                                             Pre&& pre, Stage&& stage)
{
    constexpr int N = 2 * M;
    constexpr int WPT = N / 16;             // packed words per transform
    constexpr int WTH = (F * M / 8) / NT;   // packed words per thread
    constexpr int PLO = PhaseTab<M>::PLO;
    constexpr int PHI = PhaseTab<M>::PHI;
    static_assert(WTH % 2 == 0 && M >= 128, "F kernel shape");
    static_assert(!(ONEROW && IM), "one-row groups are input-fastest");
    const long nitems = nt * ninp;
    pre();
    // ONEROW: the caller guarantees that g0 is the same in every thread and
    // that the group's items are inputs i0 .. i0 + F - 1 of one spectrum t0 (F
    // divides the input count): one division for the group, on the scalar
    // unit, instead of one per thread and use
    long t0r = 0;
    int i0r = 0;
    if constexpr (ONEROW) {
        const long gu = uniform_long(g0);
        t0r = gu / ninp;
        i0r = (int)(gu - t0r * ninp);
    }
    auto item = [&](int f, long& t, int& i) -> bool {
        if constexpr (ONEROW) {
            t = t0r;
            i = i0r + f;
            return g0 + f < nitems;
        } else {
            return fitem<IM>(g0 + f, nt, ninp, nitems, t, i);
        }
    };
    if (tid < F) {
        long t;
        int i;
        const bool v = item(tid, t, i);
        it.t[tid] = v ? (int)t : -1;
        it.i[tid] = i;
        it.amp[tid] = 0.5f * sc_uni * ((sc_inp && v) ? sc_inp[i] : 1.0f);
    }
    if (pm) {
        for (int e = tid; e < F * (PLO + PHI); e += NT) {
            const int f = e / (PLO + PHI);
            const int j = e - f * (PLO + PHI);
            long t;
            int i;
            if (!item(f, t, i)) continue;
            const fx_phase_model pmi = pm[i];
            const double tg = (double)(t_first + t);
            const double tau = pmi.delay0 + pmi.delay_rate * tg;
            // the phase in cycles, accumulated in FP64 (1/N is exact: N is a
            // power of two) and reduced to [-1/2, 1/2]
            double cyc;
            if (j < PLO) {
                cyc = (double)j * tau * (1.0 / (double)N);
            } else {
                cyc = (double)((j - PLO) * PLO) * tau * (1.0 / (double)N) +
                      (pmi.fringe0 + pmi.fringe_rate * tg) * 0.15915494309189535;
            }
            cyc -= rint(cyc);
            // then its sine and cosine in FP32 on the transcendental unit
            // (v_sin_f32 / v_cos_f32 take the angle in cycles): the reduced phase
            // in FP32 and its sincos are exact to a few 1e-8, the reference's
            // FP64-phase / FP32-sincos convention, at a tenth of an FP64
            // sincospi per table entry
            const float cf = (float)cyc;
            const float sn = __builtin_amdgcn_sinf(cf), cs = __builtin_amdgcn_cosf(cf);
            if (j < PLO) {
                // the item amplitude, as it.amp[f] holds it
                const float a = 0.5f * sc_uni * (sc_inp ? sc_inp[i] : 1.0f);
                it.ph_lo[f][j] = make_float2(cs * a, sn * a);
            } else {
                it.ph_hi[f][j - PLO] = make_float2(cs, sn);
            }
        }
    }
    uint32_t* raw = reinterpret_cast<uint32_t*>(sh);
    stage(raw);
    lds_barrier();
    {
        constexpr int S1 = M / 16;
        constexpr int BPT = (F * M / 16) / NT;
        static_assert((F * M / 16) % NT == 0, "butterflies must divide evenly");
        float2 u[BPT][16];
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            const int g = tid + b * NT;
            const int f = g / S1;
            const int j = g - f * S1;
            const uint32_t* rw = raw + f * WPT + (j >> 3);
            const int s4 = 4 * (j & 7);
#pragma unroll
            for (int r = 0; r < 16; r++) u[b][r] = tb.lut[(rw[r * (S1 >> 3)] >> s4) & 15u];
        }
        lds_barrier();
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            const int g = tid + b * NT;
            const int f = g / S1;
            const int j = g - f * S1;
            dft_reg<16>(u[b]);
            float2* o = sh + lpad(f * M + j * 16);
#pragma unroll
            for (int r = 0; r < 16; r++) o[r] = u[b][r];
        }
        lds_barrier();
    }
    hook(std::integral_constant<int, 0>{});
    fft_passes_hook<M, 16, NT, F, true, FTabs<M>::T3, 1>(sh, tb.twl, tid, tb.tw2, tb.t3, hook);
}

template <int M, int F, int NT, bool IM, typename Hook>
__device__ __forceinline__ void fstage_group(const float2* __restrict__ twg_tabs, const uint32_t* __restrict__ words,
                                             size_t in_stride, int ninp, long t_first, long nt,
                                             const fx_phase_model* __restrict__ pm, float sc_uni,
                                             const float* __restrict__ sc_inp, long g0, float2* sh, FTabs<M>& tb,
                                             FItems<M, F>& it, int tid, Hook&& hook)
{
    constexpr int WPT = M / 8;                  // packed words per transform
    constexpr int WTH = (F * M / 8) / NT;       // packed words per thread
    const int wl = WTH * tid;
    const int f = wl / WPT;
    const int wof = wl - f * WPT;
    uint2 wpre[WTH / 2];
    bool valid;
    FConstsPre<M, NT> cpre;
    // twg_tabs: also fill the LDS tables (the kernel's first group)
    fstage_group<M, F, NT, IM>(words, in_stride, ninp, t_first, nt, pm, sc_uni, sc_inp, g0, sh, tb, it, tid, hook,
                               [&]() {
        long t;
        int i;
        valid = fitem<IM>(g0 + f, nt, ninp, nt * ninp, t, i);
        // unconditional reads (an item past the chunk's end reads its input's
        // first spectrum and stages zeros), so no wait sits inside a branch
        const uint32_t* src = words + (size_t)i * in_stride + (size_t)(t_first + (valid ? t : 0)) * WPT + wof;
#pragma unroll
        for (int q = 0; q < WTH; q += 2) wpre[q / 2] = *reinterpret_cast<const uint2*>(src + q);
        if (twg_tabs) cpre.load(twg_tabs, tid);
    },
                               [&](uint32_t* raw) {
#pragma unroll
        for (int q = 0; q < WTH; q += 2)
            *reinterpret_cast<uint2*>(raw + wl + q) = valid ? wpre[q / 2] : make_uint2(0u, 0u);
        if (twg_tabs) cpre.put(tb, tid);
    });
}

template <int M, int F, int NT, bool IM>
__device__ __forceinline__ void fstage_front(const float2* __restrict__ twg, const uint32_t* __restrict__ words,
                                             size_t in_stride, int ninp, long t_first, long nt,
                                             const fx_phase_model* __restrict__ pm, float sc_uni,
                                             const float* __restrict__ sc_inp, long g0, float2* sh, FTabs<M>& tb,
                                             FItems<M, F>& it, int tid)
{
    fstage_group<M, F, NT, IM>(twg, words, in_stride, ninp, t_first, nt, pm, sc_uni, sc_inp, g0, sh, tb, it, tid,
                               [](auto) {});
}

// Channel k of item f from twice its split value: rotate by the phasor (which
// carries the item amplitude) or scale by the amplitude, then the optional
// per-channel multiplier and per-channel scale, then the engine format.
// PMO: phase model only (no multiplier, no per-channel scales), decided at
// compile time so the epilogue is straight-line code.
template <int M, typename T, int F, bool PMO>
__device__ __forceinline__ typename ChanT<T>::V fstage_finish(float2 x, const FItems<M, F>& it, int f, int k,
                                                              float2 plo, bool pm,
                                                              const float2* __restrict__ mult,
                                                              const float* __restrict__ sc_chan)
{
    constexpr int PLO = PhaseTab<M>::PLO;
    if constexpr (PMO) return to_chan<T>(cmul(x, cmul(it.ph_hi[f][k / PLO], plo)));
    if (pm)
        x = cmul(x, cmul(it.ph_hi[f][k / PLO], plo));
    else
        x = make_float2(x.x * it.amp[f], x.y * it.amp[f]);
    if (mult) x = cmul(x, mult[(size_t)it.i[f] * M + k]);
    if (sc_chan) {
        const float a = sc_chan[k];
        x.x *= a;
        x.y *= a;
    }
    return to_chan<T>(x);
}

} // namespace

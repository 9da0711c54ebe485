// fx_persist.hip -- the batch of uneven subbands in one persistent launch.
//
// k_batch_persistent walks a host-built table of work items with a
// workgroup-wide queue counter; 512-thread workgroups, two per CU:
//   PCAL  exact FP64 phase-cal fold of a range of one input's words;
//   F     F groups (2-bit unpack -> real FFT -> FP64-model rotation -> FP16),
//         8192 complex points per group, stored either straight into the
//         X-engine layout chan[c][t][i] or as natural rows nat[t][i][c];
//   CT    corner turn of natural rows into chan, 64 x 64 tiles through LDS
//         (long transforms: a group holds only 1-4 of them, so their direct
//         stores would be 4-16 byte fragments);
//   X     channels of one 64 x 64 input tile over one time slice, FP16 MFMA,
//         the two halves of the workgroup on the two channels of a pair,
//         each streaming its spectra through an LDS ring by LDS-DMA loads;
//   RED   the time slices' partial sums of a product range, in slice order;
//   TONE  the tones of one input from its fold.
// In the fused form (one launch) an item may wait for a counter of items it
// depends on (CT for the subband's F items, X for its F or CT items, RED for
// its X items, TONE for its folds); those items come earlier in the table, so
// the workgroups that took them are running and the wait ends (it is also
// bounded: a wait that outlasts PSPIN polls raises the error counter and
// skips the item).  The phased form runs the same items as four ordered
// launches (PCAL+F, CT, X, RED+TONE) without waits.
//
// The time-sliced per-core loop of the reference's dispatcher
// (TaskDispatcher.cpp:116-137) becomes this single queue of uneven work.
// Everything is deterministic: FP32 sums in a fixed order per visibility,
// exact FP64 phase-cal folds, tone sums reduced in a fixed order.
#include "fx_internal.h"
#include "fx_fstage.h"
#include "fx_xtile.h"

namespace {

constexpr int PNT = FX_PNT;       // threads per workgroup
constexpr int PFM = 8192;         // complex points per F group
constexpr int XKT = FX_PXKT;      // spectra per X stage
constexpr int XTS = 64;           // inputs per X tile side
constexpr int CTP = 65;           // corner-turn tile row pitch (words): conflict-free both ways
constexpr long PSPIN = 1L << 24;  // polls of a wait before it is abandoned (seconds of GPU time)
static_assert(PNT == 512, "the item layouts below assume 8 waves");

template <int M>
struct PGroup {
    static constexpr int F = PFM / M;   // transforms per group
};

// A pointer read from memory is a generic pointer to the compiler, and a
// generic access also waits on LDS traffic.  The subbands' arrays are all
// device memory: the kernel reads the table through this view of fx_psub,
// whose pointers are declared global, so their accesses stay global loads and
// stores.
#define FX_GLOBAL __attribute__((address_space(1)))
// plain vector types (HIP's vector classes cannot be assigned in another
// address space)
typedef unsigned int pu32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int pu32x4 __attribute__((ext_vector_type(4)));
typedef float pf32x2 __attribute__((ext_vector_type(2)));
struct PSubView {
    const uint32_t FX_GLOBAL* words;
    size_t stride;
    int nfft, ninp;
    long nt;
    const fx_phase_model FX_GLOBAL* pm;
    float sc;
    int ipad;
    uint32_t FX_GLOBAL* chan;
    uint32_t FX_GLOBAL* nat;
    float2 FX_GLOBAL* vis;
    float2 FX_GLOBAL* part;
    long ntp;
    long tslice;
    int nts;
    int pad0;
    double FX_GLOBAL* pacc;
    double FX_GLOBAL* tones;
    int nfold, nbins, ntones, offset_bin, step, rotatorlen;
    double dphi;
};
static_assert(sizeof(PSubView) == sizeof(fx_psub) && offsetof(PSubView, dphi) == offsetof(fx_psub, dphi) &&
                  offsetof(PSubView, chan) == offsetof(fx_psub, chan) &&
                  offsetof(PSubView, part) == offsetof(fx_psub, part) &&
                  offsetof(PSubView, nts) == offsetof(fx_psub, nts) && offsetof(PSubView, pacc) == offsetof(fx_psub, pacc),
              "PSubView mirrors fx_psub");

#define FX_CONST __attribute__((address_space(4)))

// The item functions are separate (not inlined) functions, so each gets the
// whole register budget of the 16 waves per CU for itself.  Their arguments
// arrive in per-lane registers: these make them uniform again, and the
// subband's view a scalar-loaded constant.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long uni(long v)
{
    const uint64_t u = (uint64_t)v;
    return (long)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(u >> 32)) << 32) |
                  (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
}
__device__ __forceinline__ const PSubView FX_CONST& psub(const fx_psub* subs, int sb)
{
    const long a = uni((long)(uintptr_t)subs);
    return reinterpret_cast<const PSubView FX_CONST*>((uintptr_t)a)[uni(sb)];
}

__host__ __device__ constexpr size_t al16(size_t n) { return (n + 15) / 16 * 16; }

// LDS of an F item: the transform tile, the tables, the items (their phasor
// tables only with a phase model)
template <int M>
__host__ __device__ constexpr size_t f_sh_bytes() { return al16((size_t)lds_complex(PFM) * sizeof(float2)); }
template <int M>
__host__ __device__ constexpr size_t f_tab_bytes() { return al16(sizeof(FTabs<M>)); }
template <int M>
__host__ __device__ constexpr size_t f_lds_bytes(bool pm)
{
    return f_sh_bytes<M>() + f_tab_bytes<M>() +
           (pm ? sizeof(FItems<M, PGroup<M>::F>) : fitems_head_bytes<M, PGroup<M>::F>());
}

// LDS of an X item: two halves x RING stages x (one or two 8 KiB slabs)
constexpr int XSLAB = XKT * XTS * 4;
template <bool DIAG>
__host__ __device__ constexpr int x_ring() { return DIAG ? 4 : 2; }
template <bool DIAG>
__host__ __device__ constexpr size_t x_lds_bytes() { return (size_t)2 * x_ring<DIAG>() * (DIAG ? 1 : 2) * XSLAB; }

template <int N>
__device__ __forceinline__ void p_wait_vmcnt()
{
    // s_waitcnt simm16 on gfx9: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | (((N >> 4) & 3) << 14));
}

// F groups [g_first, g_first + ngroups) of subband S: group g holds items
// g*F .. g*F+F-1 in the order item = t*I + i, each an unpack -> FFT ->
// rotate of one (spectrum, input).  Each thread's packed words of the next
// group load while the current group transforms.
template <int M>
__device__ __noinline__ void f_item(const float2* __restrict__ twg_, const fx_psub* subs, int sb, long g_first_,
                                    long ngroups_, unsigned char* lds)
{
    constexpr int F = PGroup<M>::F;
    constexpr int N = 2 * M;
    constexpr int PLO = PhaseTab<M>::PLO;
    constexpr int WPT = M / 8;               // packed words per transform
    constexpr int WTH = (F * M / 8) / PNT;   // packed words per thread
    static_assert(WTH == 2, "one uint2 of packed words per thread and group");
    float2* sh = reinterpret_cast<float2*>(lds);
    FTabs<M>& tb = *reinterpret_cast<FTabs<M>*>(lds + f_sh_bytes<M>());
    FItems<M, F>& it = *reinterpret_cast<FItems<M, F>*>(lds + f_sh_bytes<M>() + f_tab_bytes<M>());
    const int tid = threadIdx.x;
    const PSubView FX_CONST& S = psub(subs, sb);
    const long g_first = uni(g_first_), ngroups = uni(ngroups_);
    const float2* twg = reinterpret_cast<const float2*>(uni((long)(uintptr_t)twg_));
    const uint32_t FX_GLOBAL* const words = S.words;
    const size_t stride = S.stride;
    const int I = S.ninp;
    const long nt = S.nt;
    const fx_phase_model* const pmp = (const fx_phase_model*)S.pm;
    const float sc = S.sc;
    const int ipad = S.ipad;
    uint32_t FX_GLOBAL* const chan = S.chan;
    uint32_t FX_GLOBAL* const natr = S.nat;
    const long nitems = nt * (long)I;
    const bool pm = pmp != nullptr;
    const bool nat = (const void*)natr != nullptr;
    const size_t crow = (size_t)S.ntp * (size_t)ipad;   // one channel of chan
    auto tw_of = [&](int k) {
        if constexpr (FTabs<M>::TWL) {
            return twl_get(tb.twl, k * (FX_TW_N / N));
        } else {
            const pf32x2 v = ((const pf32x2 FX_GLOBAL*)(uintptr_t)twg)[k * (FX_TW_N / N)];
            return make_float2(v.x, v.y);
        }
    };
    auto wload = [&](long grp) -> pu32x2 {
        const int wl = WTH * tid;
        const int f = wl / WPT;
        const long g = grp * F + f;
        if (g >= nitems) return pu32x2{0u, 0u};
        const long t = g / I;
        const int i = (int)(g - t * I);
        return *reinterpret_cast<const pu32x2 FX_GLOBAL*>(words + (size_t)i * stride + (size_t)t * WPT +
                                                          (wl - f * WPT));
    };
    fstage_consts<M, PNT>(twg, tb, tid);
    pu32x2 pre = wload(g_first);
    int tl = tid;
    for (long gg = 0; gg < ngroups; gg++) {
        if (gg) {
            __syncthreads();   // the previous group's stores have read sh and it
            // tl is tid re-derived every group (an empty asm the compiler cannot
            // see through): the table lookups of a group are not hoisted out of
            // the loop into registers held across all groups
            asm volatile("" : "+v"(tl));
        }
        const long grp = g_first + gg;
        fstage_group<M, F, PNT, false>((const uint32_t*)words, stride, I, 0, nt, pmp, sc, nullptr, grp * F, sh, tb,
                                       it, tl, [](auto) {}, []() {},
                                       [&](uint32_t* raw) {
            *reinterpret_cast<pu32x2*>(raw + WTH * tl) = pre;
            if (gg + 1 < ngroups) pre = wload(grp + 1);
        });
        auto fin = [&](float2 x, int f, int k) {
            return fstage_finish<M, __half2, F, false>(x, it, f, k, pm ? it.ph_lo[f][k % PLO] : make_float2(1.f, 0.f),
                                                      pm, nullptr, nullptr);
        };
        if (!nat) {
            // Channel pairs (kp, M - kp), items fastest: neighbouring lanes
            // store neighbouring inputs of one spectrum.  F divides PNT, so a
            // thread keeps one item f and steps through its channels by
            // STEP = PNT / F (8 pairs per thread for every M): the item's
            // column, the store addresses and the split twiddle W_N^kp
            // advance by constant steps (the twiddle by one complex multiply).
            constexpr int STEP = PNT / F;
            const int f = tl & (F - 1);
            const int k0 = tl / F;
            if (it.t[f] >= 0) {
                const size_t col = (size_t)it.t[f] * ipad + it.i[f];
                uint32_t FX_GLOBAL* pk = chan + col + (size_t)k0 * crow;
                uint32_t FX_GLOBAL* pm2 = chan + col + (size_t)(M - k0) * crow;
                const size_t dstep = (size_t)STEP * crow;
                float2 w = tw_of(k0);
                const float2 ws = tw_of(STEP);
#pragma unroll 2
                for (int kp = k0; kp < M / 2; kp += STEP) {
                    float2 xk, xm;
                    rsplit2_tw<M>(sh, f * M, kp, w, xk, xm);
                    *pk = fin(xk, f, kp);
                    if (kp) *pm2 = fin(xm, f, M - kp);   // kp = 0: M - kp is the dropped Nyquist bin
                    w = cmul(w, ws);
                    pk += dstep;
                    pm2 -= dstep;
                }
            }
            if (tl < F && it.t[tl] >= 0) {
                float2 xk, xm;
                rsplit2_tw<M>(sh, tl * M, M / 2, tw_of(M / 2), xk, xm);
                chan[(size_t)(M / 2) * crow + (size_t)it.t[tl] * ipad + it.i[tl]] = fin(xk, tl, M / 2);
            }
        } else if constexpr (M >= 1024) {
            // natural rows (the host sets them from 2048-point transforms):
            // neighbouring lanes store neighbouring channels, KPT pairs per
            // thread and item, the split twiddle advancing by W_N^PNT
            constexpr int KPT = (M / 2) / PNT;
            const float2 ws = tw_of(PNT);
            for (int f = 0; f < F; f++) {
                if (it.t[f] < 0) continue;
                uint32_t FX_GLOBAL* row = natr + ((size_t)it.t[f] * ipad + it.i[f]) * (size_t)M;
                float2 w = tw_of(tl);
#pragma unroll
                for (int kk = 0; kk < KPT; kk++) {
                    const int kp = tl + kk * PNT;
                    float2 xk, xm;
                    rsplit2_tw<M>(sh, f * M, kp, w, xk, xm);
                    row[kp] = fin(xk, f, kp);
                    if (kp) row[M - kp] = fin(xm, f, M - kp);
                    w = cmul(w, ws);
                }
            }
            if (tl < F && it.t[tl] >= 0) {
                float2 xk, xm;
                rsplit2_tw<M>(sh, tl * M, M / 2, tw_of(M / 2), xk, xm);
                natr[((size_t)it.t[tl] * ipad + it.i[tl]) * (size_t)M + M / 2] = fin(xk, tl, M / 2);
            }
        }
    }
}

// Corner turn of the natural rows of spectra [t0, t0 + ntt) into chan, for
// channel block cb (64 channels) and input block ib (64 inputs): per spectrum
// a 64 x 64 tile read as rows of 64 channels, written as rows of 64 inputs,
// the next spectrum's rows loading while the current tile turns.
__device__ __noinline__ void ct_item(const fx_psub* subs, int sb, long t0_, long ntt_, int blk_, unsigned char* lds)
{
    uint32_t* tile = reinterpret_cast<uint32_t*>(lds);   // [64 channels][CTP]
    const int tid = threadIdx.x;
    const PSubView FX_CONST& S = psub(subs, sb);
    const long t0 = uni(t0_), ntt = uni(ntt_);
    const int blk = uni(blk_);
    const int I = S.ninp;
    const int ipad = S.ipad;
    const uint32_t FX_GLOBAL* const natr = S.nat;
    uint32_t FX_GLOBAL* const chan = S.chan;
    const int nib = (I + 63) / 64;
    const int cb = blk / nib, ib = blk - cb * nib;
    const long C = S.nfft / 2;
    const int i0 = ib * 64;
    const long cc0 = (long)cb * 64;
    const size_t crow = (size_t)S.ntp * (size_t)ipad;
    const long t1 = t0 + ntt < S.nt ? t0 + ntt : S.nt;
    // read map: input rows rr and rr + 32, channels rs .. rs + 3
    const int rr = tid >> 4, rs = (tid & 15) * 4;
    // write map: channel wc, inputs wi .. wi + 7
    const int wc = tid >> 3, wi = (tid & 7) * 8;
    pu32x4 v[2];
    auto load = [&](long t) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int i = i0 + rr + 32 * k;
            v[k] = (t < t1 && i < I)
                       ? *reinterpret_cast<const pu32x4 FX_GLOBAL*>(natr + ((size_t)t * ipad + i) * (size_t)C + cc0 + rs)
                       : pu32x4{0u, 0u, 0u, 0u};
        }
    };
    load(t0);
    for (long t = t0; t < t1; t++) {
        if (t > t0) __syncthreads();   // the previous spectrum's tile has been read
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int r = rr + 32 * k;
            tile[(rs + 0) * CTP + r] = v[k].x;
            tile[(rs + 1) * CTP + r] = v[k].y;
            tile[(rs + 2) * CTP + r] = v[k].z;
            tile[(rs + 3) * CTP + r] = v[k].w;
        }
        __syncthreads();
        load(t + 1);
        uint32_t o[8];
#pragma unroll
        for (int e = 0; e < 8; e++) o[e] = tile[wc * CTP + wi + e];
        uint32_t FX_GLOBAL* dst = chan + (size_t)(cc0 + wc) * crow + (size_t)t * ipad + i0 + wi;
        if (i0 + wi + 8 <= I) {
            reinterpret_cast<pu32x4 FX_GLOBAL*>(dst)[0] = pu32x4{o[0], o[1], o[2], o[3]};
            reinterpret_cast<pu32x4 FX_GLOBAL*>(dst)[1] = pu32x4{o[4], o[5], o[6], o[7]};
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (i0 + wi + e < I) dst[e] = o[e];
        }
    }
}

// X work of channels [c0, c0 + nc) of tile (ti, tj) over time slice `slice`.
// Waves 0-3 take the even channels of the range, waves 4-7 the odd ones;
// each half streams its KT-spectrum slabs (one for a diagonal tile, the row
// and column blocks otherwise) through its own ring of LDS stages filled by
// LDS-DMA loads (global_load_lds_dwordx4: lane l's 16 bytes land at the
// wave's LDS base + 16 l), RING - 1 stages in flight, continuing from one
// channel to the next.  Wave (qi, qj) of a half accumulates the 32 x 32
// block (qi, qj) of the tile (the strictly-lower block of a diagonal tile
// idles).  A slab row holds 64 inputs from row0: rows past the subband's
// inputs read padding or the next spectrum's inputs and only reach products
// that are not stored.  With one slice the visibilities go to vis with the F
// stage's amplitude scale removed, else the raw sums to part[slice].
template <bool DIAG>
__device__ __noinline__ void x_item(const fx_psub* subs, int sb, int c0_, int nc_, int ti_, int tj_, int slice_,
                                    unsigned char* lds)
{
    constexpr int RING = x_ring<DIAG>();
    constexpr int NSLAB = DIAG ? 1 : 2;
    constexpr int STAGE = NSLAB * XSLAB;
    constexpr int NPER = STAGE / 1024 / 4;   // 1 KiB loads per wave per stage
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const int half = w >> 2, hw = w & 3;
    const int qi = hw >> 1, qj = hw & 1;
    const bool active = !(DIAG && qi > qj);
    unsigned char* ring = lds + half * (RING * STAGE);
    const PSubView FX_CONST& S = psub(subs, sb);
    const int c0 = uni(c0_), nc = uni(nc_), ti = uni(ti_), tj = uni(tj_), slice = uni(slice_);
    const int I = S.ninp;
    const int C = S.nfft / 2;
    const int ipad = S.ipad;
    const uint32_t FX_GLOBAL* const chan = S.chan;
    const int nts = S.nts;
    const float d = S.sc * S.sc;
    pf32x2 FX_GLOBAL* const outb = reinterpret_cast<pf32x2 FX_GLOBAL*>(nts == 1 ? S.vis : S.part) +
                                   (nts == 1 ? 0 : (size_t)slice * ((size_t)I * (I + 1) / 2) * C);
    const int t0 = slice * (int)S.tslice;
    const int t1 = t0 + (int)S.tslice < (int)S.ntp ? t0 + (int)S.tslice : (int)S.ntp;
    const int nst = (t1 - t0) / XKT;   // >= 1: the host sizes the slices
    const int G = ((nc + 1) / 2) * nst;
    const size_t crow = (size_t)S.ntp * (size_t)ipad;
    // this lane's part of every 1 KiB load: spectrum row 4 b8 + lrow of the
    // stage, inputs lcol .. lcol + 3 of the slab's 64
    const int loff = (lane >> 4) * ipad + (lane & 15) * 4;
    // the next stage to load: channel pair ip, stage ist
    int ip = 0, ist = 0, nissued = 0;
    auto issue = [&]() {
        int c = c0 + 2 * ip + half;
        if (c >= c0 + nc) c = c0 + 2 * ip;   // odd count: the idle half reloads its partner's channel
        const uint32_t FX_GLOBAL* src = chan + (size_t)c * crow + (size_t)(t0 + ist * XKT) * ipad + loff;
        unsigned char* dst = ring + (nissued % RING) * STAGE;
#pragma unroll
        for (int k = 0; k < NPER; k++) {
            const int blk = hw * NPER + k;   // 1 KiB block of the stage
            const int s = blk >> 3, b8 = blk & 7;
            const uint32_t FX_GLOBAL* a = src + (4 * b8) * ipad + (s ? tj : ti) * XTS;
            __builtin_amdgcn_global_load_lds((FX_GLOBAL void*)a,
                                             (__attribute__((address_space(3))) void*)(dst + blk * 1024), 16, 0, 0);
        }
        nissued++;
        if (++ist == nst) {
            ist = 0;
            ip++;
        }
    };
    fx_float16 dre = fx_float16{}, dim = fx_float16{};
    const int r32 = lane & 31, h = lane >> 5;
    p_wait_vmcnt<0>();   // nothing of an earlier item is outstanding in the counts below
    for (int g = 0; g < RING - 1 && g < G; g++) issue();
    bool drain = false;
    int p = 0, st = 0;   // the stage consumed: channel pair p, stage st
    for (int g = 0; g < G; g++) {
        // this wave's loads of stage g are done once at most the later
        // stages' loads are outstanding (after a channel's stores: none);
        // the barrier makes every wave's part of the stage visible and frees
        // the slot of stage g - 1
        const int ahead = G - 1 - g;
        if (drain) p_wait_vmcnt<0>();
        else if (ahead >= RING - 2) p_wait_vmcnt<(RING - 2) * NPER>();
        else if (ahead == 1) p_wait_vmcnt<NPER>();
        else p_wait_vmcnt<0>();
        drain = false;
        // a plain barrier: __syncthreads() would also wait for the later
        // stages' loads (its release fence drains vmcnt)
        asm volatile("s_barrier" ::: "memory");
        if (nissued < G) issue();
        if (active) {
            const uint32_t* sa = reinterpret_cast<const uint32_t*>(ring + (g % RING) * STAGE) + 4 * h * XTS +
                                 qi * 32 + r32;
            const uint32_t* sb = reinterpret_cast<const uint32_t*>(ring + (g % RING) * STAGE + (DIAG ? 0 : XSLAB)) +
                                 4 * h * XTS + qj * 32 + r32;
#pragma unroll 2
            for (int ks = 0; ks < XKT / 8; ks++) {
                fx_int4v ia, ib, ir;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ia[q] = (int)sa[(ks * 8 + q) * XTS];
                    const uint32_t v = sb[(ks * 8 + q) * XTS];
                    ib[q] = (int)v;
                    ir[q] = (int)rot_h2(v);
                }
                const fx_half8 fa = __builtin_bit_cast(fx_half8, ia);
                dre = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa, __builtin_bit_cast(fx_half8, ib), dre, 0, 0, 0);
                dim = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa, __builtin_bit_cast(fx_half8, ir), dim, 0, 0, 0);
            }
        }
        if (++st == nst) {
            const int c = c0 + 2 * p + half;
            if (active && c < c0 + nc) {
                const int j = tj * XTS + qj * 32 + r32;
                const int ib0 = ti * XTS + qi * 32 + 4 * h;
                pf32x2 FX_GLOBAL* out = outb + c;
                const float dd = nts == 1 ? d : 1.f;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int i = ib0 + 8 * (r >> 2) + (r & 3);
                    if (i <= j && j < I) {
                        const int pp = i * I - i * (i - 1) / 2 + (j - i);
                        out[(size_t)pp * C] = pf32x2{dre[r] / dd, dim[r] / dd};
                    }
                }
            }
            dre = fx_float16{};
            dim = fx_float16{};
            drain = true;
            st = 0;
            p++;
        }
    }
}

// X work of channels [c0, c0 + nc) of a subband of 65 to 128 inputs over time
// slice `slice`, as one 128 x 128 tile: the workgroup streams one channel's
// 128-input slabs through a ring of LDS stages (LDS-DMA loads, as x_item),
// and its 8 waves share the tile's 10 upper 32 x 32 blocks (waves 0 and 1 take
// two).  Reading each spectrum's inputs once instead of as the three 64-input
// tiles' four slabs halves the X item's HBM traffic.
__device__ __noinline__ void x_item128(const fx_psub* subs, int sb, int c0_, int nc_, int slice_, unsigned char* lds)
{
    constexpr int RING = 4;
    constexpr int STAGE = XKT * 128 * 4;      // 16 KiB
    constexpr int NPER = STAGE / 1024 / 8;    // 1 KiB loads per wave per stage
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const PSubView FX_CONST& S = psub(subs, sb);
    const int c0 = uni(c0_), nc = uni(nc_), slice = uni(slice_);
    const int I = S.ninp;
    const int C = S.nfft / 2;
    const int ipad = S.ipad;
    const uint32_t FX_GLOBAL* const chan = S.chan;
    const int nts = S.nts;
    const float d = S.sc * S.sc;
    pf32x2 FX_GLOBAL* const outb = reinterpret_cast<pf32x2 FX_GLOBAL*>(nts == 1 ? S.vis : S.part) +
                                   (nts == 1 ? 0 : (size_t)slice * ((size_t)I * (I + 1) / 2) * C);
    const int t0 = slice * (int)S.tslice;
    const int t1 = t0 + (int)S.tslice < (int)S.ntp ? t0 + (int)S.tslice : (int)S.ntp;
    const int nst = (t1 - t0) / XKT;   // >= 1: the host sizes the slices
    const int G = nc * nst;
    const size_t crow = (size_t)S.ntp * (size_t)ipad;
    // blocks (a, b), a <= b, of the 4 x 4 grid: wave w takes block w and
    // waves 0, 1 also blocks 8, 9
    constexpr int BA[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    constexpr int BB[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    const int a0 = BA[w], b0 = BB[w];
    const bool two = w < 2;
    const int a1 = two ? BA[8 + w] : 0, b1 = two ? BB[8 + w] : 0;
    // this lane's part of every 1 KiB load: spectrum row 2 blk + (lane >> 5),
    // inputs (lane & 31) * 4 .. + 3
    const int loff = (lane >> 5) * ipad + (lane & 31) * 4;
    int ic = 0, ist = 0, nissued = 0;
    auto issue = [&]() {
        const uint32_t FX_GLOBAL* src = chan + (size_t)(c0 + ic) * crow + (size_t)(t0 + ist * XKT) * ipad + loff;
        unsigned char* dst = lds + (nissued % RING) * STAGE;
#pragma unroll
        for (int k = 0; k < NPER; k++) {
            const int blk = w * NPER + k;
            __builtin_amdgcn_global_load_lds((FX_GLOBAL void*)(src + (2 * blk) * ipad),
                                             (__attribute__((address_space(3))) void*)(dst + blk * 1024), 16, 0, 0);
        }
        nissued++;
        if (++ist == nst) {
            ist = 0;
            ic++;
        }
    };
    fx_float16 re0 = fx_float16{}, im0 = fx_float16{}, re1 = fx_float16{}, im1 = fx_float16{};
    const int r32 = lane & 31, h = lane >> 5;
    p_wait_vmcnt<0>();   // nothing of an earlier item is outstanding in the counts below
    for (int g = 0; g < RING - 1 && g < G; g++) issue();
    bool drain = false;
    int c = c0, st = 0;
    for (int g = 0; g < G; g++) {
        const int ahead = G - 1 - g;
        if (drain) p_wait_vmcnt<0>();
        else if (ahead >= RING - 2) p_wait_vmcnt<(RING - 2) * NPER>();
        else if (ahead == 1) p_wait_vmcnt<NPER>();
        else p_wait_vmcnt<0>();
        drain = false;
        asm volatile("s_barrier" ::: "memory");
        if (nissued < G) issue();
        const uint32_t* sl = reinterpret_cast<const uint32_t*>(lds + (g % RING) * STAGE) + 4 * h * 128 + r32;
        auto mac = [&](int a, int b, fx_float16& dre, fx_float16& dim) {
#pragma unroll 2
            for (int ks = 0; ks < XKT / 8; ks++) {
                fx_int4v ia, ib, ir;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ia[q] = (int)sl[(ks * 8 + q) * 128 + a * 32];
                    const uint32_t v = sl[(ks * 8 + q) * 128 + b * 32];
                    ib[q] = (int)v;
                    ir[q] = (int)rot_h2(v);
                }
                const fx_half8 fa = __builtin_bit_cast(fx_half8, ia);
                dre = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa, __builtin_bit_cast(fx_half8, ib), dre, 0, 0, 0);
                dim = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa, __builtin_bit_cast(fx_half8, ir), dim, 0, 0, 0);
            }
        };
        mac(a0, b0, re0, im0);
        if (two) mac(a1, b1, re1, im1);
        if (++st == nst) {
            const float dd = nts == 1 ? d : 1.f;
            auto put = [&](int a, int b, const fx_float16& dre, const fx_float16& dim) {
                // the lane re-derived (an empty asm): the 16 product offsets
                // are computed here, not hoisted out of the stage loop and
                // held (or spilled) across it
                int ln = lane;
                asm volatile("" : "+v"(ln));
                const int j = b * 32 + (ln & 31);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int i = a * 32 + 8 * (r >> 2) + 4 * (ln >> 5) + (r & 3);
                    if (i <= j && j < I) {
                        const int pp = i * I - i * (i - 1) / 2 + (j - i);
                        outb[(size_t)pp * C + c] = pf32x2{dre[r] / dd, dim[r] / dd};
                    }
                }
            };
            put(a0, b0, re0, im0);
            if (two) put(a1, b1, re1, im1);
            re0 = im0 = re1 = im1 = fx_float16{};
            drain = true;
            st = 0;
            c++;
        }
    }
}

// vis[p][c] of products [p0, p0 + np): the slices' raw sums added in slice
// order, the F stage's amplitude scale removed
__device__ __noinline__ void red_item(const fx_psub* subs, int sb, long p0_, long np_)
{
    const PSubView FX_CONST& S = psub(subs, sb);
    const long p0 = uni(p0_), np = uni(np_);
    const long C = S.nfft / 2;
    const long P = (long)S.ninp * (S.ninp + 1) / 2;
    const float d = S.sc * S.sc;
    const int nts = S.nts;
    const pf32x2 FX_GLOBAL* const part = reinterpret_cast<const pf32x2 FX_GLOBAL*>(S.part);
    pf32x2 FX_GLOBAL* const vis = reinterpret_cast<pf32x2 FX_GLOBAL*>(S.vis);
    const size_t n = (size_t)np * C, base = (size_t)p0 * C, sstr = (size_t)P * C;
    for (size_t e = threadIdx.x; e < n; e += PNT) {
        float re = 0.f, im = 0.f;
        for (int s = 0; s < nts; s++) {
            const pf32x2 v = part[s * sstr + base + e];
            re += v.x;
            im += v.y;
        }
        vis[base + e] = pf32x2{re / d, im / d};
    }
}

// Phase-cal fold of words [w_first, w_first + nwords) of input y (w_first a
// multiple of nfold/16): pacc[y][(16 w + j) mod nfold] += LUT(code).  A
// thread takes the words of one residue w mod nfold/16, so its 16 bins stay
// the same and it counts codes per bin in registers (Pcal16: four positions
// per 32-bit add; bin sum H A + B); the FP64 sums are exact, so their order
// does not matter.
__device__ __noinline__ void pcal_item(const fx_psub* subs, int sb, int y_, long w_first_, long nwords_,
                                       unsigned char* lds)
{
    double* lacc = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const PSubView FX_CONST& S = psub(subs, sb);
    const int y = uni(y_);
    const long w_first = uni(w_first_), nwords = uni(nwords_);
    const int nfold = S.nfold;
    const int nb = nfold / 16;
    for (int b = tid; b < nfold; b += PNT) lacc[b] = 0.0;
    __syncthreads();
    const uint32_t FX_GLOBAL* words = S.words + (size_t)y * S.stride;
    const long nw = (long)S.nt * S.nfft / 16;
    const long w_end = w_first + nwords < nw ? w_first + nwords : nw;
    const int lanes = PNT / nb;   // threads per residue
    if (tid < lanes * nb) {
        const int r = tid % nb;
        Pcal16 cnt;
        cnt.init();
        for (long w = w_first + r + (long)nb * (tid / nb); w < w_end; w += (long)nb * lanes) cnt.add(words[w]);
        cnt.flush();
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (cnt.A[j] | cnt.B[j])
                atomicAdd(&lacc[16 * r + j], (double)FX_LUT_HI * (double)cnt.A[j] + (double)cnt.B[j]);
    }
    __syncthreads();
    double FX_GLOBAL* acc = S.pacc + (size_t)y * (size_t)nfold;
    for (int b = tid; b < nfold; b += PNT)
        if (lacc[b] != 0.0) atomicAdd((double*)&acc[b], lacc[b]);
}

// The tones of input y from its fold (the tone extraction of k_pcal_final:
// tone t at DFT bin offset_bin + t*step of the nbins-point fold, the
// rotator folded in for the Shifting extractor), partial sums reduced in a
// fixed order
__device__ __noinline__ void tone_item(const fx_psub* subs, int sb, int y_, unsigned char* lds)
{
    double(*red)[PNT] = reinterpret_cast<double(*)[PNT]>(lds);
    const int tid = threadIdx.x;
    const PSubView FX_CONST& S = psub(subs, sb);
    const int y = uni(y_);
    const double FX_GLOBAL* acc = S.pacc + (size_t)y * (size_t)S.nfold;
    double FX_GLOBAL* tones = S.tones + (size_t)y * 2 * (size_t)S.ntones;
    for (int t = 0; t < S.ntones; t++) {
        const long k = (long)S.offset_bin + (long)t * S.step;
        double sr = 0.0, si = 0.0;
        if (k < S.nbins) {
            for (int r = tid; r < S.nfold; r += PNT) {
                const double v = acc[r];
                if (v == 0.0) continue;
                const long e = ((long)(r % S.nbins) * k) % S.nbins;
                double ang = -6.283185307179586 * (double)e / (double)S.nbins;
                if (S.rotatorlen) ang += S.dphi * (double)(r % S.rotatorlen);
                double sn, cs;
                sincos(ang, &sn, &cs);
                sr += v * cs;
                si += v * sn;
            }
        }
        if (t) __syncthreads();   // the previous tone's result has been read
        red[0][tid] = sr;
        red[1][tid] = si;
        __syncthreads();
        for (int hh = PNT / 2; hh > 0; hh >>= 1) {
            if (tid < hh) {
                red[0][tid] += red[0][tid + hh];
                red[1][tid] += red[1][tid + hh];
            }
            __syncthreads();
        }
        if (tid == 0) {
            tones[2 * t] = k < S.nbins ? red[0][0] / (double)S.nbins : 0.0;
            tones[2 * t + 1] = k < S.nbins ? red[1][0] / (double)S.nbins : 0.0;
        }
    }
}

template <typename Fn>
__device__ __forceinline__ void with_m(int nfft, Fn&& fn)
{
    switch (nfft) {
    case 256: fn(std::integral_constant<int, 128>{}); break;
    case 512: fn(std::integral_constant<int, 256>{}); break;
    case 1024: fn(std::integral_constant<int, 512>{}); break;
    case 2048: fn(std::integral_constant<int, 1024>{}); break;
    case 4096: fn(std::integral_constant<int, 2048>{}); break;
    case 8192: fn(std::integral_constant<int, 4096>{}); break;
    case 16384: fn(std::integral_constant<int, 8192>{}); break;
    default: break;
    }
}

__global__ __launch_bounds__(PNT, 4) void k_batch_persistent(const float2* __restrict__ twg,
                                                          const fx_psub* __restrict__ subs,
                                                          const fx_pitem* __restrict__ items, int i0, int i1,
                                                          int* __restrict__ ctrs, int head, int wait)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_it, s_ok;
    const int tid = threadIdx.x;
    // the last wait this workgroup completed: a later item waiting for the
    // same counter and at most that count needs neither the polls nor another
    // acquire (the data it reads was already visible after the first)
    int seen_ctr = -1, seen_val = 0;
    for (;;) {
        if (tid == 0) s_it = i0 + atomicAdd(&ctrs[head], 1);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(s_it);
        if (k >= i1) break;   // the queue is empty: every workgroup gets here
        const fx_pitem FX_CONST& item = reinterpret_cast<const fx_pitem FX_CONST*>((uintptr_t)items)[k];
        const int type = item.type, sb = item.sb;
        const long ia = item.a, ib = item.b;
        const int ic = item.c, id = item.d;
        const int wait_ctr = item.wait_ctr, wait_val = item.wait_val, done_ctr = item.done_ctr;
        if (wait && wait_ctr >= 0 && !(wait_ctr == seen_ctr && wait_val <= seen_val)) {
            if (tid == 0) {
                // poll without ordering; one acquire once the count is reached
                int ok = 1;
                for (long spin = 0;
                     __hip_atomic_load(&ctrs[wait_ctr], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < wait_val;
                     spin++) {
                    if (spin >= PSPIN ||
                        __hip_atomic_load(&ctrs[FX_PCTR_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        ok = 0;
                        atomicOr(&ctrs[FX_PCTR_ERR], 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(16);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                s_ok = ok;
            }
            __syncthreads();       // the other waves read after thread 0's acquire
            if (!s_ok) continue;   // abandoned (the error counter is set)
            seen_ctr = wait_ctr;
            seen_val = wait_val;
        }
        const PSubView FX_CONST& S = reinterpret_cast<const PSubView FX_CONST*>((uintptr_t)subs)[sb];
        switch (type) {
        case FX_PI_F:
            with_m(S.nfft, [&](auto m) { f_item<decltype(m)::value>(twg, subs, sb, ia, ib, lds); });
            break;
        case FX_PI_CT:
            ct_item(subs, sb, ia, ib, ic, lds);
            break;
        case FX_PI_X: {
            if (item.xt == 128) {
                x_item128(subs, sb, (int)ia, (int)ib, id, lds);
                break;
            }
            int ti, tj;
            tile_coords(ic, (S.ninp + XTS - 1) / XTS, ti, tj);
            if (ti == tj)
                x_item<true>(subs, sb, (int)ia, (int)ib, ti, tj, id, lds);
            else
                x_item<false>(subs, sb, (int)ia, (int)ib, ti, tj, id, lds);
            break;
        }
        case FX_PI_RED:
            red_item(subs, sb, ia, ib);
            break;
        case FX_PI_PCAL:
            pcal_item(subs, sb, (int)ia, ib, ic, lds);
            break;
        case FX_PI_TONE:
            tone_item(subs, sb, (int)ia, lds);
            break;
        default:
            break;
        }
        if (wait && done_ctr >= 0) {
            // every wave's stores are done (the barrier's workgroup release);
            // thread 0's agent-scope release makes them visible device-wide
            __syncthreads();
            if (tid == 0) __hip_atomic_fetch_add(&ctrs[done_ctr], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();           // LDS and s_it are free for the next item
    }
}

template <int M>
size_t f_lds_of(bool pm) { return f_lds_bytes<M>(pm); }

} // namespace

int fx_persist_group_items(int nfft) { return PFM / (nfft / 2); }

size_t fx_persist_lds(int type, int arg, bool pm)
{
    switch (type) {
    case FX_PI_F:
        switch (arg) {
        case 256: return f_lds_of<128>(pm);
        case 512: return f_lds_of<256>(pm);
        case 1024: return f_lds_of<512>(pm);
        case 2048: return f_lds_of<1024>(pm);
        case 4096: return f_lds_of<2048>(pm);
        case 8192: return f_lds_of<4096>(pm);
        case 16384: return f_lds_of<8192>(pm);
        default: return 0;
        }
    case FX_PI_X: return arg ? x_lds_bytes<true>() : x_lds_bytes<false>();   // x_item128: 4 x 16 KiB, the same
    case FX_PI_CT: return (size_t)64 * CTP * sizeof(uint32_t);
    case FX_PI_PCAL: return (size_t)arg * sizeof(double);
    case FX_PI_TONE: return (size_t)2 * PNT * sizeof(double);
    default: return 0;
    }
}

// dynamic LDS above the default 64 KiB needs the kernel attribute raised
static int persist_attr(size_t lds_bytes)
{
    static size_t set = 0;
    if (lds_bytes <= set) return FX_OK;
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_persistent),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
        hipGetLastError();
        fx_set_error("persistent batch kernel: %zu bytes of dynamic LDS refused: %s", lds_bytes, hipGetErrorString(e));
        return FX_ERR_HIP;
    }
    set = lds_bytes;
    return FX_OK;
}

int fxk_batch_persistent_blocks(size_t lds_bytes)
{
    if (persist_attr(lds_bytes) != FX_OK) return 0;
    int dev = 0, ncu = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        hipGetLastError();
        fx_set_error("persistent batch kernel: no device attributes");
        return 0;
    }
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void*>(&k_batch_persistent), PNT, lds_bytes);
    if (e != hipSuccess || occ < 1) {
        hipGetLastError();
        // from the LDS alone: 160 KiB per CU
        occ = (int)std::max<size_t>(1, (160u * 1024u) / std::max<size_t>(lds_bytes + 64, 1));
    }
    return std::min(occ, 2) * std::max(1, ncu);
}

int fxk_batch_persistent(hipStream_t s, const float2* tw, const fx_psub* subs, const fx_pitem* items, int i0, int i1,
                         int* ctrs, int head, int nblocks, int wait, size_t lds_bytes)
{
    if (i1 <= i0) return FX_OK;
    if (head < 0 || head >= FX_PCTR_ERR || nblocks < 1 || lds_bytes > 160 * 1024) return FX_ERR_ARG;
    if (persist_attr(lds_bytes) != FX_OK) return FX_ERR_HIP;
    hipLaunchKernelGGL(k_batch_persistent, dim3((unsigned)nblocks), dim3(PNT), lds_bytes, s, tw, subs, items, i0, i1,
                       ctrs, head, wait);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

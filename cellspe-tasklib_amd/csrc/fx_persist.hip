// fx_persist.hip -- the batch of uneven subbands in one persistent launch.
//
// k_batch_persistent walks a host-built table of work items with a
// workgroup-wide queue counter: phase-cal folds, F groups (2-bit unpack ->
// real FFT -> FP64-model rotation -> FP16 [c][t][i] store), X tiles (the
// FP16 MFMA engine over a subband's whole integration, writing its
// visibilities) and phase-cal tone extractions, for subbands of any FFT length
// and input count in one table.  An X item waits until every F group of its
// subband is done and a tone item until every fold of its subband is done;
// the items it waits for come earlier in the table, so the workgroups that
// took them are running and the wait ends (it is also bounded: a wait that
// outlasts PSPIN polls raises the error counter and skips the item).  The
// same table can run as phased launches instead (folds and F groups, then X
// tiles, then tones), which order the items by launch and skip the waits.
//
// The time-sliced per-core loop of the reference's dispatcher
// (TaskDispatcher.cpp:116-137) becomes this single queue of uneven work.
// Everything is deterministic: FP32 sums in a fixed order per visibility,
// exact FP64 phase-cal folds, tone sums reduced in a fixed order.
#include "fx_internal.h"
#include "fx_fstage.h"
#include "fx_xtile.h"

namespace {

constexpr int PNT = XNT;          // threads per workgroup
constexpr int PFM = 4096;         // complex points per F group (one transform from M = 8192 up)
constexpr long PSPIN = 1L << 24;  // polls of a wait before it is abandoned (seconds of GPU time)

template <int M>
struct PGroup {
    static constexpr int FM = M > PFM ? M : PFM;
    static constexpr int F = FM / M;   // transforms per group
};

// A pointer read from memory is a generic pointer to the compiler, and a
// generic access also waits on LDS traffic.  The subbands' arrays are all
// device memory: the kernel reads the table through this view of fx_psub,
// whose pointers are declared global, so their accesses stay global loads and
// stores.
#define FX_GLOBAL __attribute__((address_space(1)))
struct PSubView {
    const uint32_t FX_GLOBAL* words;
    size_t stride;
    int nfft, ninp;
    long nt;
    const fx_phase_model FX_GLOBAL* pm;
    float sc;
    uint32_t FX_GLOBAL* chan;
    float2 FX_GLOBAL* vis;
    double FX_GLOBAL* pacc;
    double FX_GLOBAL* tones;
    int nfold, nbins, ntones, offset_bin, step, rotatorlen;
    double dphi;
};
static_assert(sizeof(PSubView) == sizeof(fx_psub) && offsetof(PSubView, dphi) == offsetof(fx_psub, dphi) &&
                  offsetof(PSubView, chan) == offsetof(fx_psub, chan) && offsetof(PSubView, pacc) == offsetof(fx_psub, pacc),
              "PSubView mirrors fx_psub");

__host__ __device__ constexpr size_t al16(size_t n) { return (n + 15) / 16 * 16; }

// LDS of an F item: the transform tile, the tables, the items' phasor tables
template <int M>
__host__ __device__ constexpr size_t f_sh_bytes() { return al16((size_t)lds_complex(PGroup<M>::FM) * sizeof(float2)); }
template <int M>
__host__ __device__ constexpr size_t f_tab_bytes() { return al16(sizeof(FTabs<M>)); }
template <int M>
__host__ __device__ constexpr size_t f_lds_bytes()
{
    return f_sh_bytes<M>() + f_tab_bytes<M>() + sizeof(FItems<M, PGroup<M>::F>);
}

// F groups [g_first, g_first + ngroups) of subband S: group g holds items
// g*F .. g*F+F-1 in the order item = t*I + i, each an unpack -> FFT ->
// rotate of one (spectrum, input); every channel goes to chan[c][t][i].
template <int M>
__device__ void f_item(const float2* __restrict__ twg, const PSubView& S, long g_first, long ngroups,
                       unsigned char* lds)
{
    constexpr int F = PGroup<M>::F;
    constexpr int N = 2 * M;
    constexpr int PLO = PhaseTab<M>::PLO;
    float2* sh = reinterpret_cast<float2*>(lds);
    FTabs<M>& tb = *reinterpret_cast<FTabs<M>*>(lds + f_sh_bytes<M>());
    FItems<M, F>& it = *reinterpret_cast<FItems<M, F>*>(lds + f_sh_bytes<M>() + f_tab_bytes<M>());
    const int tid = threadIdx.x;
    const int I = S.ninp;
    const size_t row = (size_t)S.nt * (size_t)I;   // one channel of the whole integration
    const bool pm = (const void*)S.pm != nullptr;
    auto tw_of = [&](int k) {
        if constexpr (FTabs<M>::TWL)
            return twl_get(tb.twl, k * (FX_TW_N / N));
        else
            return twg[k * (FX_TW_N / N)];
    };
    fstage_consts<M, PNT>(twg, tb, tid);
    int tl = tid;
    for (long gg = 0; gg < ngroups; gg++) {
        if (gg) {
            __syncthreads();   // the previous group's stores have read sh and it
            // tl is tid re-derived every group (an empty asm the compiler cannot
            // see through): the table lookups of a group are not hoisted out of
            // the loop into registers held across all groups
            asm volatile("" : "+v"(tl));
        }
        fstage_group<M, F, PNT, false>((const uint32_t*)S.words, S.stride, I, 0, S.nt, (const fx_phase_model*)S.pm,
                                       S.sc, nullptr, (g_first + gg) * F, sh, tb,
                                       it, tl, [](auto) {});
        // channel pairs (kp, M - kp), items fastest: neighbouring lanes store
        // neighbouring inputs of one spectrum
#pragma unroll 1
        for (int q = tl; q < F * (M / 2); q += PNT) {
            const int f = q % F;
            const int kp = q / F;
            if (it.t[f] < 0) continue;
            float2 xk, xm;
            rsplit2_tw<M>(sh, f * M, kp, tw_of(kp), xk, xm);
            uint32_t FX_GLOBAL* col = S.chan + (size_t)it.t[f] * I + it.i[f];
            col[(size_t)kp * row] = fstage_finish<M, __half2, F, false>(
                xk, it, f, kp, pm ? it.ph_lo[f][kp % PLO] : make_float2(1.f, 0.f), pm, nullptr, nullptr);
            if (kp) {   // kp = 0: M - kp is the dropped Nyquist bin
                const int k2 = M - kp;
                col[(size_t)k2 * row] = fstage_finish<M, __half2, F, false>(
                    xm, it, f, k2, pm ? it.ph_lo[f][k2 % PLO] : make_float2(1.f, 0.f), pm, nullptr, nullptr);
            }
        }
        if (tl < F && it.t[tl] >= 0) {
            float2 xk, xm;
            rsplit2_tw<M>(sh, tl * M, M / 2, tw_of(M / 2), xk, xm);
            S.chan[(size_t)(M / 2) * row + (size_t)it.t[tl] * I + it.i[tl]] = fstage_finish<M, __half2, F, false>(
                xk, it, tl, M / 2, pm ? it.ph_lo[tl][(M / 2) % PLO] : make_float2(1.f, 0.f), pm, nullptr, nullptr);
        }
    }
}

// X tile `tile` (TSZ x TSZ inputs) of channels [c0, c0 + nc) over the
// subband's whole integration; visibilities vis[p][c] with the F stage's
// amplitude scale removed (p = i*I - i(i-1)/2 + (j-i), i <= j)
template <int TSZ>
__device__ void x_item(const PSubView& S, long c0, long nc, int tile, unsigned char* lds)
{
    constexpr int KT = XTraits<FX_PREC_F16>::KT;
    constexpr int NBW = TSZ / 64;
    uint32_t(*slab)[KT][TSZ] = reinterpret_cast<uint32_t(*)[KT][TSZ]>(lds);
    const int I = S.ninp;
    const int nb = (I + TSZ - 1) / TSZ;
    int ti, tj;
    tile_coords(tile, nb, ti, tj);
    const long C = S.nfft / 2;
    const float d = S.sc * S.sc;
    for (long c = c0; c < c0 + nc; c++) {
        if (c > c0) __syncthreads();   // the previous channel's last stage has been read
        xengine_tile<FX_PREC_F16, TSZ>((const uint32_t*)(S.chan + (size_t)c * (size_t)S.nt * (size_t)I), S.nt, I, ti,
                                       tj, slab,
                                       [&](const auto& dre, const auto& dim, int w, int qi, int qj, bool diag,
                                           int lane) {
#pragma unroll
            for (int a = 0; a < NBW; a++)
#pragma unroll
                for (int b = 0; b < NBW; b++) {
                    if (diag && qi == qj && a > b) continue;
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int i = ti * TSZ + (NBW * qi + a) * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        const int j = tj * TSZ + (NBW * qj + b) * 32 + (lane & 31);
                        if (i <= j && j < I) {
                            const long p = (long)i * I - (long)i * (i - 1) / 2 + (j - i);
                            ((float2*)S.vis)[(size_t)p * C + c] = make_float2(dre[a][b][r] / d, dim[a][b][r] / d);
                        }
                    }
                }
        });
    }
}

// Phase-cal fold of words [w_first, w_first + nwords) of input y (w_first a
// multiple of nfold/16): pacc[y][(16 w + j) mod nfold] += LUT(code).  A
// thread takes the words of one residue w mod nfold/16, so its 16 bins stay
// the same and it counts codes per bin in registers (A = n(+H) - n(-H),
// B = n(+1) - n(-1), bin sum H A + B); the FP64 sums are exact, so their
// order does not matter.
__device__ void pcal_item(const PSubView& S, int y, long w_first, long nwords, unsigned char* lds)
{
    double* lacc = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
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
        int A[16], B[16];
#pragma unroll
        for (int j = 0; j < 16; j++) A[j] = B[j] = 0;
        for (long w = w_first + r + (long)nb * (tid / nb); w < w_end; w += (long)nb * lanes) {
            const uint32_t v = words[w];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t c = (v >> (2 * j)) & 3u;
                A[j] += (int)(c == 3u) - (int)(c == 0u);
                B[j] += (int)(c == 2u) - (int)(c == 1u);
            }
        }
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (A[j] | B[j]) atomicAdd(&lacc[16 * r + j], (double)FX_LUT_HI * (double)A[j] + (double)B[j]);
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
__device__ void tone_item(const PSubView& S, int y, unsigned char* lds)
{
    double(*red)[PNT] = reinterpret_cast<double(*)[PNT]>(lds);
    const int tid = threadIdx.x;
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
        for (int h = PNT / 2; h > 0; h >>= 1) {
            if (tid < h) {
                red[0][tid] += red[0][tid + h];
                red[1][tid] += red[1][tid + h];
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

__global__ __launch_bounds__(PNT, 2) void k_batch_persistent(const float2* __restrict__ twg,
                                                          const fx_psub* __restrict__ subs,
                                                          const fx_pitem* __restrict__ items, int i0, int i1,
                                                          int* __restrict__ ctrs, int head, int wait)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_it, s_ok;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_it = i0 + atomicAdd(&ctrs[head], 1);
        __syncthreads();
        const int k = s_it;
        if (k >= i1) break;   // the queue is empty: every workgroup gets here
        const fx_pitem item = items[k];
        const PSubView& S = reinterpret_cast<const PSubView*>(subs)[item.sb];
        if (wait && item.wait_ctr >= 0) {
            if (tid == 0) {
                int ok = 1;
                for (long spin = 0; __hip_atomic_load(&ctrs[item.wait_ctr], __ATOMIC_ACQUIRE,
                                                      __HIP_MEMORY_SCOPE_AGENT) < item.wait_val;
                     spin++) {
                    if (spin >= PSPIN ||
                        __hip_atomic_load(&ctrs[FX_PCTR_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        ok = 0;
                        atomicOr(&ctrs[FX_PCTR_ERR], 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(16);
                }
                s_ok = ok;
            }
            __syncthreads();
            if (!s_ok) continue;   // abandoned (the error counter is set)
            __threadfence();       // every wave sees what the awaited items stored
        }
        switch (item.type) {
        case FX_PI_F:
            with_m(S.nfft, [&](auto m) { f_item<decltype(m)::value>(twg, S, item.a, item.b, lds); });
            break;
        case FX_PI_X:
            x_item<64>(S, item.a, item.b, item.c, lds);
            break;
        case FX_PI_PCAL:
            pcal_item(S, (int)item.a, item.b, item.c, lds);
            break;
        case FX_PI_TONE:
            tone_item(S, (int)item.a, lds);
            break;
        default:
            break;
        }
        if (item.done_ctr >= 0) {
            __threadfence();       // this wave's stores reach device scope
            __syncthreads();
            if (tid == 0) __hip_atomic_fetch_add(&ctrs[item.done_ctr], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();           // LDS and s_it are free for the next item
    }
}

template <int M>
size_t f_lds_of() { return f_lds_bytes<M>(); }

} // namespace

int fx_persist_group_items(int nfft)
{
    const int M = nfft / 2;
    return (M > PFM ? M : PFM) / M;
}

size_t fx_persist_lds(int type, int arg)
{
    switch (type) {
    case FX_PI_F:
        switch (arg) {
        case 256: return f_lds_of<128>();
        case 512: return f_lds_of<256>();
        case 1024: return f_lds_of<512>();
        case 2048: return f_lds_of<1024>();
        case 4096: return f_lds_of<2048>();
        case 8192: return f_lds_of<4096>();
        case 16384: return f_lds_of<8192>();
        default: return 0;
        }
    case FX_PI_X: return (size_t)2 * XTraits<FX_PREC_F16>::KT * (size_t)arg * sizeof(uint32_t);
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
    return occ * std::max(1, ncu);
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

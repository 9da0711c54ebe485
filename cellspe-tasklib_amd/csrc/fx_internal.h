// fx_internal.h -- shared definitions of the gfx950 F-X engine (HIP).
//
// Device-side helpers (complex arithmetic, LDS padding, small register DFTs,
// the LDS Stockham FFT) and the host-side launcher declarations every
// translation unit of libcellspe_fx.so shares.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stddef.h>
#include <algorithm>
#include <type_traits>
#include <vector>

#include "cellspe_fx.h"

// ---------------------------------------------------------------------------
// Error handling
// ---------------------------------------------------------------------------
#define FX_HIP_TRY(expr)                                                        \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            fx_set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,          \
                         hipGetErrorString(_e));                                \
            return FX_ERR_HIP;                                                  \
        }                                                                       \
    } while (0)

#define FX_REQUIRE(cond, code, ...)                                             \
    do {                                                                        \
        if (!(cond)) {                                                          \
            fx_set_error(__VA_ARGS__);                                          \
            return (code);                                                      \
        }                                                                       \
    } while (0)

void fx_set_error(const char* fmt, ...);

// ---------------------------------------------------------------------------
// Constants
// ---------------------------------------------------------------------------
// Master twiddle table: T[e] = exp(-2 pi i e / FX_TW_N), derived in FP64 and
// rounded once to FP32.  Every FFT of complex length M <= FX_TW_N/2 reads its
// twiddles W_M^e = T[e * FX_TW_N / M] and the real-FFT split twiddles
// W_N^k = T[k * FX_TW_N / N] from it.
#define FX_TW_N 32768
#define FX_MAX_N 32768      // largest real FFT length
#define FX_LUT_HI 3.3359f   // mark5access OPTIMAL_2BIT_HIGH

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
// Complex product with its two roundings fixed (one product, then one fused
// multiply-add per component), so every kernel that inlines it rounds the
// same way whatever the surrounding code lets the compiler contract.
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.y, b.x, a.x * b.y));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// LDS index padding: one spare complex every 16 keeps the strided Stockham
// accesses off a single bank, and makes every stride that is a multiple of 16
// a constant offset: lpad(i + 16 m) = lpad(i) + 17 m.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_complex(int n) { return n + (n >> 4); }
// padded distance of a stride that is a multiple of 16 (or 1 inside a block)
__host__ __device__ constexpr int lstride(int s) { return s + (s >> 4); }

// Offset-binary 2-bit level table {-H, -1, +1, +H}.
__device__ __forceinline__ float lut2(uint32_t code)
{
    // branch-free select; exact constants
    float m = (code == 0u || code == 3u) ? FX_LUT_HI : 1.0f;
    return (code & 2u) ? m : -m;
}

// Exact per-position code counts of packed 2-bit words, for the phase-cal
// folds: the 16 codes of a word sit at positions j = 0..15 (bits 2j, 2j+1).
// Byte b of the lane counters n3[s], n2[s], n1[s] counts codes 3, 2, 1 at
// position 4b + s, so one 32-bit add counts four positions at once; every
// 255 words flush() folds them into A[j] = n(+H) - n(-H) = 2 n3 + n2 + n1 - n
// and B[j] = n(+1) - n(-1) = n2 - n1 (n: words counted).  A level-table sum
// is then H A[j] + B[j], the same integers as counting code by code.
struct Pcal16 {
    int A[16], B[16];
    uint32_t n3[4], n2[4], n1[4];
    int batch;
    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int j = 0; j < 16; j++) A[j] = B[j] = 0;
#pragma unroll
        for (int s = 0; s < 4; s++) n3[s] = n2[s] = n1[s] = 0u;
        batch = 0;
    }
    __device__ __forceinline__ void flush()
    {
#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int c3 = (int)((n3[s] >> (8 * b)) & 255u), c2 = (int)((n2[s] >> (8 * b)) & 255u),
                          c1 = (int)((n1[s] >> (8 * b)) & 255u);
                A[4 * b + s] += 2 * c3 + c2 + c1 - batch;
                B[4 * b + s] += c2 - c1;
            }
            n3[s] = n2[s] = n1[s] = 0u;
        }
        batch = 0;
    }
    __device__ __forceinline__ void add(uint32_t v)
    {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const uint32_t x = (v >> (2 * s)) & 0x03030303u;
            const uint32_t lo = x & 0x01010101u, hi = (x >> 1) & 0x01010101u, p3 = hi & lo;
            n3[s] += p3;
            n2[s] += hi ^ p3;
            n1[s] += lo ^ p3;
        }
        if (++batch == 255) flush();
    }
    // the first `live` codes of a word only (the stream's last, partial word)
    __device__ __forceinline__ void add_partial(uint32_t v, int live)
    {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t c = (v >> (2 * j)) & 3u;
            const int on = j < live;
            A[j] += on * ((int)(c == 3u) - (int)(c == 0u));
            B[j] += on * ((int)(c == 2u) - (int)(c == 1u));
        }
    }
};

// cos/sin(2 pi k/16), k = 0..7 (forward sign: -sin)
__device__ constexpr float kW16re[8] = {1.0f, 0.9238795325f, 0.7071067812f, 0.3826834324f,
                                       0.0f, -0.3826834324f, -0.7071067812f, -0.9238795325f};
__device__ constexpr float kW16im[8] = {-0.0f, -0.3826834324f, -0.7071067812f, -0.9238795325f,
                                       -1.0f, -0.9238795325f, -0.7071067812f, -0.3826834324f};

// Workgroup barrier for data exchanged through LDS only: the wave waits for
// its own LDS operations, not for its outstanding global loads and stores (a
// __syncthreads() also drains those, so global stores issued before it could
// not overlap the work after it).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// In-register forward DFT of R points (R = 1,2,4,8,16), decimation in time.
template <int R>
__device__ __forceinline__ void dft_reg(float2* u)
{
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        float2 a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    } else {
        constexpr int H = R / 2;
        float2 e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; i++) {
            e[i] = u[2 * i];
            o[i] = u[2 * i + 1];
        }
        dft_reg<H>(e);
        dft_reg<H>(o);
#pragma unroll
        for (int k = 0; k < H; k++) {
            float2 t;
            constexpr int S = 16 / R;
            if (k == 0) {
                t = o[k];
            } else if (2 * k == H) {
                t = cmul_mi(o[k]);
            } else {
                t = cmul(o[k], make_float2(kW16re[k * S], kW16im[k * S]));
            }
            u[k] = cadd(e[k], t);
            u[k + H] = csub(e[k], t);
        }
    }
}

// Twiddles come from a two-level LDS table: W^e = T_hi[e >> 8] * T_lo[e & 255]
// with T_lo[j] = W^j (j < 256) and T_hi[j] = W^{256 j} (j < 128), W =
// exp(-2 pi i / FX_TW_N).  384 entries (3 KiB) replace gathers from the 256 KiB
// global table; the product costs one complex multiply (about 1 ulp).
#define FX_TWL 384
__device__ __forceinline__ void twl_load(float2* twl, const float2* __restrict__ tw, int tid, int nt)
{
    for (int j = tid; j < FX_TWL; j += nt) twl[j] = (j < 256) ? tw[j] : tw[(j - 256) << 8];
}
__device__ __forceinline__ float2 twl_get(const float2* twl, int e)
{
    float2 lo = twl[e & 255];
    int hi = e >> 8;
    return hi ? cmul(twl[256 + hi], lo) : lo;
}

// XCD-aware block index: the hardware hands block b to XCD b % 8, so logical
// block (b % 8) * per + b / 8 gives every XCD a contiguous range of logical
// blocks (neighbouring tiles share that XCD's L2).  The grid is padded to a
// multiple of 8; logical blocks >= nlogical return.
__host__ __device__ inline long xcd_grid(long nlogical) { return ((nlogical + 7) / 8) * 8; }
__device__ __forceinline__ long xcd_block(long nlogical)
{
    long per = (nlogical + 7) / 8;
    long b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}

// Twiddles w^1 .. w^15 of one radix-16 butterfly from three table lookups
// (w, w^4, w^8) and twelve products, at most two products deep from a looked-up
// value: p[r] = W^{r e}, W = exp(-2 pi i / FX_TW_N).
__device__ __forceinline__ void tw16_from(float2 w1, float2 w4, float2 w8, float2* p)
{
    p[1] = w1;
    p[2] = cmul(w1, w1);
    p[3] = cmul(p[2], w1);
    p[4] = w4;
    p[5] = cmul(w4, w1);
    p[6] = cmul(w4, p[2]);
    p[7] = cmul(w4, p[3]);
    p[8] = w8;
    p[12] = cmul(w8, w4);
#pragma unroll
    for (int r = 1; r < 4; r++) {
        p[8 + r] = cmul(w8, p[r]);
        p[12 + r] = cmul(p[12], p[r]);
    }
}
__device__ __forceinline__ void tw16_powers(const float2* twl, int e, float2* p)
{
    tw16_from(twl_get(twl, e & (FX_TW_N - 1)), twl_get(twl, (4 * e) & (FX_TW_N - 1)),
              twl_get(twl, (8 * e) & (FX_TW_N - 1)), p);
}

// One radix-R Stockham pass over F complex FFTs of length M held in LDS
// (padded indexing), NT threads, current sub-transform span P.
// Butterfly j of FFT f reads x[j + r M/R], multiplies by W_{PR}^{r k}
// (k = j mod P), does the R-point DFT and stores y[(j-k) R + k + r P].
// Strides that are multiples of 16 (and the 16 consecutive stores of the
// first pass) are constant LDS offsets from one padded base.
// With T2, the P = 16 radix-16 pass reads its twiddles W_256^{r k} from the
// table tw2[r*16 + k] instead of deriving them; with T3 the P = 256 radix-16
// pass reads w = W_4096^k, w^4 and w^8 from t3[k], t3[256 + k], t3[512 + k]
// (lanes read consecutive entries: no LDS bank conflicts).
template <int M, int R, int P, int NT, int F, bool T2 = false, bool T3 = false>
__device__ __forceinline__ void stockham_pass(float2* sh, const float2* __restrict__ tw, int tid,
                                              const float2* __restrict__ tw2 = nullptr,
                                              const float2* __restrict__ t3 = nullptr)
{
    constexpr int NB = F * M / R;
    constexpr int BPT = NB / NT;
    static_assert(NB % NT == 0, "butterflies must divide evenly");
    constexpr int S = M / R;
    constexpr bool CL = (S % 16) == 0;
    constexpr bool CS = (P % 16) == 0 || (P == 1 && R == 16);
    float2 u[BPT][R];
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int g = tid + b * NT;
        int f = g / S;
        int j = g - f * S;
        if constexpr (CL) {
            const int base = lpad(f * M + j);
#pragma unroll
            for (int r = 0; r < R; r++) u[b][r] = sh[base + r * lstride(S)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) u[b][r] = sh[lpad(f * M + j + r * S)];
        }
    }
    lds_barrier();
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        int g = tid + b * NT;
        int f = g / S;
        int j = g - f * S;
        int k = j & (P - 1);
        if constexpr (P > 1) {
            if constexpr (R == 16 && P == 16 && T2) {
#pragma unroll
                for (int r = 1; r < R; r++) u[b][r] = cmul(u[b][r], tw2[r * 16 + k]);
            } else if constexpr (R == 16 && P == 256 && T3) {
                float2 p[16];
                tw16_from(t3[k], t3[256 + k], t3[512 + k], p);
#pragma unroll
                for (int r = 1; r < R; r++) u[b][r] = cmul(u[b][r], p[r]);
            } else if constexpr (R == 16) {
                float2 p[16];
                tw16_powers(tw, k * (FX_TW_N / (P * R)), p);
#pragma unroll
                for (int r = 1; r < R; r++) u[b][r] = cmul(u[b][r], p[r]);
            } else {
#pragma unroll
                for (int r = 1; r < R; r++) u[b][r] = cmul(u[b][r], twl_get(tw, (r * k) * (FX_TW_N / (P * R))));
            }
        }
        dft_reg<R>(u[b]);
        int base = f * M + (j - k) * R + k;
        if constexpr (CS) {
            const int bl = lpad(base);
#pragma unroll
            for (int r = 0; r < R; r++) sh[bl + r * lstride(P)] = u[b][r];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) sh[lpad(base + r * P)] = u[b][r];
        }
    }
    lds_barrier();
}

template <int M, int P, int NT, int F, bool T2 = false>
__device__ __forceinline__ void fft_passes(float2* sh, const float2* __restrict__ tw, int tid,
                                           const float2* __restrict__ tw2 = nullptr)
{
    if constexpr (P < M) {
        constexpr int REM = M / P;
        constexpr int R = REM >= 16 ? 16 : REM;
        stockham_pass<M, R, P, NT, F, T2>(sh, tw, tid, tw2);
        fft_passes<M, P * R, NT, F, T2>(sh, tw, tid, tw2);
    }
}

// Number of Stockham passes from span P to M.
template <int M, int P>
__host__ __device__ constexpr int fft_npasses()
{
    if constexpr (P < M)
        return 1 + fft_npasses<M, P * (M / P >= 16 ? 16 : M / P)>();
    else
        return 0;
}

// fft_passes with hook(std::integral_constant<int, p>) called after pass p
// (counted from 0 at span P0), e.g. to issue deferred global stores between
// passes.
template <int M, int P, int NT, int F, bool T2, bool T3, int PI = 0, typename Hook>
__device__ __forceinline__ void fft_passes_hook(float2* sh, const float2* __restrict__ tw, int tid,
                                                const float2* __restrict__ tw2, const float2* __restrict__ t3,
                                                Hook&& hook)
{
    if constexpr (P < M) {
        constexpr int REM = M / P;
        constexpr int R = REM >= 16 ? 16 : REM;
        stockham_pass<M, R, P, NT, F, T2, T3>(sh, tw, tid, tw2, t3);
        hook(std::integral_constant<int, PI>{});
        fft_passes_hook<M, P * R, NT, F, T2, T3, PI + 1>(sh, tw, tid, tw2, t3, hook);
    }
}

// Real-FFT split of pair (k, M-k) of one FFT held in LDS:
//   E = (Z_k + conj Z_{M-k})/2, O = -i (Z_k - conj Z_{M-k})/2,
//   X_k = E + W_N^k O,  X_{M-k} = conj(E - W_N^k O),  N = 2M.
// For k = 0 only X_0 = Re Z_0 + Im Z_0 is a channel (Nyquist dropped).
// same split with the twiddle w = W_N^k supplied by the caller
template <int M>
__device__ __forceinline__ void rsplit_tw(const float2* sh, int base, int k, float2 w, float2& xk, float2& xmk)
{
    float2 a = sh[lpad(base + k)];
    float2 b = sh[lpad(base + ((M - k) & (M - 1)))];
    float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    float2 o = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
    float2 wo = cmul(w, o);
    xk = cadd(e, wo);
    xmk = make_float2(e.x - wo.x, -(e.y - wo.y));
}

// the same split without its factor 1/2: returns 2 X_k and 2 X_{M-k}
template <int M>
__device__ __forceinline__ void rsplit2_tw(const float2* sh, int base, int k, float2 w, float2& xk, float2& xmk)
{
    float2 a = sh[lpad(base + k)];
    float2 b = sh[lpad(base + ((M - k) & (M - 1)))];
    float2 e = make_float2(a.x + b.x, a.y - b.y);
    float2 o = make_float2(a.y + b.y, b.x - a.x);
    float2 wo = cmul(w, o);
    xk = cadd(e, wo);
    xmk = make_float2(e.x - wo.x, wo.y - e.y);
}

template <int M>
__device__ __forceinline__ void rsplit(const float2* sh, int base, int k, const float2* __restrict__ tw,
                                      float2& xk, float2& xmk)
{
    rsplit_tw<M>(sh, base, k, twl_get(tw, k * (FX_TW_N / (2 * M))), xk, xmk);
}

// ---------------------------------------------------------------------------
// Launchers (defined next to their kernels)
// ---------------------------------------------------------------------------
int fxk_unpack2b(hipStream_t s, const uint32_t* words, size_t ntime, int nchan, float* out);
int fxk_synth_station(hipStream_t s, uint64_t seed, int station, double rho, long delay,
                      double tone_amp, double tone_freq, double thresh, size_t nsamp,
                      uint32_t* words);
int fxk_synth_multichannel(hipStream_t s, uint64_t seed, int nchan, const double* tone_amp,
                           const double* tone_freq, double thresh, size_t ntime, uint32_t* words);
int fxk_synth_channelized(hipStream_t s, uint64_t seed, int nchan, long ntime, int ninp,
                          double rho, double thresh, int imajor, uint8_t* codes);

int fxk_twiddle_init(hipStream_t s, float2* tw);
// inverse: the unnormalized backward transform times scale
int fxk_fft_c2c(hipStream_t s, const float2* tw, const float2* in, float2* out, int m, long batch,
                int inverse = 0, float scale = 1.0f);
int fxk_fft_r2c(hipStream_t s, const float2* tw, const float* in, float2* out, int n, long batch,
                int keep_nyquist);
// window table of kind FX_WIN_* and n points, FP64 closed forms rounded to FP32
int fxk_window(hipStream_t s, int kind, int n, float* w);
// Launch shape of fx_spectrometer_ex: spectra, their sample stride, the FFT
// hop, the partial-sum groups per (spectrum, channel) and the bytes of
// partial spectra the launch needs (fxk_spectrometer's partial buffer).
struct fx_spec_plan {
    long nspec, spec_stride, hop;
    int groups;
    size_t part_bytes;
};
// validates the arguments of fx_spectrometer_ex and fills the plan
int fx_spec_make_plan(size_t ntime, int nchan, int n, int overlap, int integrate_ffts, int window, int continuous,
                      fx_spec_plan* plan);
int fxk_spectrometer(hipStream_t s, const float2* tw, const float* window, const uint32_t* words,
                     long hop, long spec_stride, int nchan, int n, int integrate_ffts, int nspec, float* partial,
                     int groups, int nyq, float* out);
// scale_chan: per-channel scales, or nullptr when every channel is scaled by
// scale_uni; scale_inp: per-input scales or nullptr
int fxk_fstage(hipStream_t s, const float2* tw, const uint32_t* words, size_t in_stride_words,
               int ninp, int n, long t_first, long nt, const fx_phase_model* pm,
               const float2* mult, const float* scale_chan, float scale_uni, const float* scale_inp,
               int out_type, void* nat, void* out);
// F kernel writing the engine layout itself (out[c][col0 + g] with rows of
// ostride columns; FP8 out[c][i*ostride_i + col0 + t']) in one pass, for real
// lengths n = 4096, 8192, 16384 with F = 32768/n items per workgroup;
// fxk_fstage_direct_ok says whether a launch of these shapes is supported
// nwg: workgroups of the persistent form (one per CU; FP32/FP16 out, at most
// 256 inputs; ctr its 8 block counters), or 0 for one block per workgroup;
// pair: the persistent form takes FP16 blocks in pairs where the inputs allow
int fxk_fstage_direct(hipStream_t s, const float2* tw, const uint32_t* words, size_t in_stride_words,
                      int ninp, int n, long t_first, long nt, const fx_phase_model* pm,
                      const float2* mult, const float* scale_chan, float scale_uni, const float* scale_inp,
                      int out_type, void* out, long ostride, long col0, long ostride_i, unsigned* ctr, int nwg,
                      bool pair);
bool fxk_fstage_direct_ok(int n, int ninp, long nt, int out_type, long ostride, long col0, long ostride_i);
// corner turn of one sub-chunk of F-stage rows nat[g][c] (nitems items of ns
// spectra) into the chunk layout: FP32/FP16 out[c][col0 + g] with rows of
// ostride items, FP8 out[c][i*ostride_i + col0 + t'] (fxk_fstage with
// out == nullptr stops after the F kernel)
int fxk_cornerturn(hipStream_t s, int out_type, const void* nat, void* out, long nitems, int nchan,
                   long ostride, long col0, int ns, long ostride_i);
// FP32/FP16 engine on [c][t][i], ts x ts input tiles (ts = 128 or 64); ring:
// the LDS-ring engines may run (whole 128-input tiles and whole stages),
// else the register-prefetch engine runs for every shape; prog (null: none)
// holds fxk_xengine_prog_words() words for the lockstep groups of the
// large-input ring engine (cleared by a fill kernel before each launch)
int fxk_xengine(hipStream_t s, int in_type, int ts, const void* X, int nchan, long nt, int ninp,
                float* acc, bool ring, unsigned* prog);
int fxk_xengine_prog_words();
// name of the kernel fxk_xengine launches for these arguments ("" if none)
// (lock: the large-input ring engine may run in lockstep groups, as with a
// non-null prog above)
const char* fxk_xengine_name(int in_type, int ts, const void* X, int nchan, long nt, int ninp, bool ring,
                             bool lock);
// FX_XENGINE_LOCK=1 in the environment turns the lockstep groups on by default
bool fxk_xengine_lock_default();
// FX_XENGINE_RING=0 in the environment turns the ring engines off by default
bool fxk_xengine_ring_default();
// 8-bit X-engine on input-major channelized data X[c][i][t]: FX_PREC_FP8
// (e4m3 re|im<<8 pairs, one accumulator set) or FX_PREC_FP8_2BIT (2-bit
// codes re|im<<2, three accumulator sets of the exact decomposition) or
// FX_PREC_FP4_2BIT (the same decomposition with e2m1 operands)
int fxk_xengine8(hipStream_t s, int prec, const void* X, int nchan, long nt, int ninp, float* acc);
// 2-bit channelized codes [c][t][i] or [c][i][t] -> [c][t][i] FP32/FP16 samples
int fxk_codes_to_chan(hipStream_t s, const uint8_t* codes, int in_imajor, int nchan, long nt, int ninp,
                      int out_prec, void* out);
// engine image size for ts x ts tiles (the 8-bit engines use ts = 128)
size_t fxk_xengine_acc_floats(int nchan, int ninp, int nsets, int ts);
// exact_coef: the four integer coefficients of the exact 2-bit combination
// (three accumulator sets), or null for a plain sum of the sets
int fxk_xengine_dump_pij(hipStream_t s, const float* acc, int nchan, int ninp, int ts, int nsets,
                         const long long* exact_coef, const float* scale_chan, const float* scale_inp,
                         const int2* pij, int accumulate, float2* vis);
int fxk_mfma_probe(hipStream_t s, int which, float* out);
// phase-cal of ninp inputs in one launch: input y reads words + y*in_stride_words
// and owns acc[y*nfold ..] and tones[y*2*ntones ..]
// wvalid (may be null): one byte per packed word of input 0, 0 = blanked (skipped)
int fxk_pcal_fold(hipStream_t s, const uint32_t* words, size_t nsamp, size_t position,
                  int nfold, double* acc, int ninp = 1, size_t in_stride_words = 0,
                  const uint8_t* wvalid = nullptr);
int fxk_pcal_final(hipStream_t s, const double* acc, int nfold, int nbins, int ntones,
                   int offset_bin, int step, int rotatorlen, double dphi, double* tones, int ninp = 1);
int fxk_complex_mac(hipStream_t s, const float2* a, const float2* b, float2* acc, long n, int conj);
int fxk_complex_mul(hipStream_t s, const float2* a, const float2* b, float2* out, long n);
int fxk_sincos(hipStream_t s, const float* arg, float2* out, long n, int normalized);
int fxk_sincos_gen(hipStream_t s, double start, double inc, float2* out, long n);
int fxk_copy(hipStream_t s, const void* src, void* dst, size_t bytes);
int fxk_complex_add(hipStream_t s, const float2* a, const float2* b, float2* out, long n);
int fxk_iotest(hipStream_t s, float* out);
// fill n 32-bit words with value (the graph-capturable clear)
int fxk_fill32(hipStream_t s, void* p, uint32_t value, size_t n);
// place nframes VDIF/Mark5B frames of raw (one workgroup per frame) into
// out rows by (second, frame number, thread); valid[(row*nslots + slot)*2]
// gets the valid 32-bit word range of the slot's payload; counts: placed,
// placed but blank, slot out of range, unmapped thread
// windowed 2-bit FFT inputs of one channel, frames g0..g0+nframes-1 (see
// k_spec_frames); wvalid (may be null): one byte per packed word, 0 = blanked
int fxk_spec_frames(hipStream_t s, const uint32_t* words, const uint8_t* wvalid, int nchan, int ch, long hop,
                    long spec_stride, int K, int n, long g0, long nframes, const float* win, float* out);
// cross spectra with optional per-word validity of either stream (see fx_spectrometer_xpol)
int fx_spec_xpol_impl(fx_ctx* ctx, const uint32_t* d_words_a, const uint8_t* d_valid_a, int nchan_a, int chan_a,
                      const uint32_t* d_words_b, const uint8_t* d_valid_b, int nchan_b, int chan_b, size_t ntime,
                      int n, int overlap, int integrate_ffts, int window, int continuous, fx_cf32* d_out,
                      long* nspec_out);
// per-word validity of nslots placed payloads of pwords words from their
// [start, end) ranges (valid[2 slot], valid[2 slot + 1]); flags[slot * pwords + j]
int fxk_valid_words(hipStream_t s, const uint32_t* valid, long nslots, int pwords, uint8_t* flags);
// *any = 1 if one of the n flags is 0 (else left as is)
int fxk_any_zero(hipStream_t s, const uint8_t* flags, size_t n, int* any);
// per spectrum and bin: scale * sum over K FFTs of A conj(B) (DC, Nyquist: re*re),
// the FFTs split over `groups` partial sums (part: groups * nspec * c1 float2)
int fxk_xpol_groups(int c1, int K, long nspec);
int fxk_xpol_acc(hipStream_t s, const float2* A, const float2* B, int c1, int K, long nspec, float scale,
                 float2* part, int groups, float2* out);
// phase-cal of nspec TaskCore spectra of L samples at d_words (1 channel),
// spectrum indices counted from spec0; fold: grown device scratch; kuse
// (< 0: K): FFTs of each spectrum taken (a trailing partial spectrum has
// fewer than K)
int fx_spec_pcal(fx_ctx* ctx, const uint32_t* d_words, long nspec, long L, long spec0, int N, int ov, int K, int R,
                 int T, double dphi, void*& fold, size_t& fold_cap, std::vector<fx_cf64>& out,
                 const uint8_t* d_wvalid = nullptr, int kuse = -1);
int fxk_frames_place(hipStream_t s, int format, const uint8_t* raw, long nframes, int framebytes,
                     int payload_offset, const int* thread_row, int map_size, int nrows, long sec0, long frame0,
                     int fps,
                     long nslots, uint32_t* out, size_t out_stride_words, uint32_t* valid,
                     unsigned long long* counts, bool exact);
// amplitude |z|*scale and phase atan2(im, re) of n complex values (either output may be null)
int fxk_amp_phase(hipStream_t s, const float2* z, long n, float scale, float* amp, float* phase);
// per input: the m5pcal adjacent-tone delay fit (ns) of ntones tones at
// f0 + t*df Hz, tones[y*ntones + t]
int fxk_pcal_delay(hipStream_t s, const double* tones, int ninp, int ntones, double f0_hz, double df_hz,
                   double band_center_hz, double band_valid_hz, double* delay_ns);
int fxk_difx_post(hipStream_t s, const float2* fft, int nfft, int band0, int n, const float2* mult,
                  float2* out, float2* conj_out, float2* acc);

// ---------------------------------------------------------------------------
// Persistent batch kernel (fx_persist.hip): one launch walks a table of work
// items over many subbands -- phase-cal folds, F groups, corner turns, X
// tiles, slice reductions, phase-cal tones -- with 512-thread workgroups.
//
// Subband layouts: chan[c][t][i] half2, ntp rows of ipad inputs per channel
// (ipad = ninp rounded up to 4, ntp = nt rounded up to 32, the pad never
// written and zero); nat[t][i][c] (rows t*ipad + i of nfft/2 channels) when
// the F items write natural rows and corner-turn items build chan; part[s][p][c]
// the raw sums of time slice s when the X items are split over nts slices.
struct fx_psub {
    const uint32_t* words;      // input 0; input i at words + i*stride
    size_t stride;
    int nfft, ninp;
    long nt;                    // spectra integrated
    const fx_phase_model* pm;   // device array of ninp models, or null
    float sc;                   // uniform amplitude scale before the FP16 cast
    int ipad;                   // inputs per chan / nat row
    uint32_t* chan;             // [c][ntp][ipad] half2 samples
    uint32_t* nat;              // [nt][ipad][nfft/2] natural rows, or null (F items write chan)
    float2* vis;                // [p][c] visibilities (scale removed)
    float2* part;               // [nts][p][c] raw slice sums (nts > 1), else null
    long ntp;                   // spectra rows of chan (multiple of 32)
    long tslice;                // spectra per time slice (multiple of 32)
    int nts;                    // time slices
    int pad0;
    double* pacc;               // phase-cal fold, ninp x nfold (null: no phase-cal)
    double* tones;              // ninp x ntones complex
    int nfold, nbins, ntones, offset_bin, step, rotatorlen;
    double dphi;
};
enum { FX_PI_PCAL = 0, FX_PI_F = 1, FX_PI_X = 2, FX_PI_TONE = 3, FX_PI_CT = 4, FX_PI_RED = 5 };
struct fx_pitem {
    int type, sb;
    long a, b;      // F: first group, groups; X: first channel, channels; CT: first spectrum, spectra;
                    // PCAL: input, first word; TONE: input; RED: first product, products
    int c;          // X: tile index; CT: channel block * (input blocks) + input block; PCAL: words per item
    int d;          // X: time slice
    int wait_ctr;   // counter this item waits for (-1: none)
    int wait_val;   // value that counter must reach
    int done_ctr;   // counter incremented when the item is done (-1: none)
    int xt;         // X: tile width, 64 or 128 (one tile of up to 128 inputs)
};
// counters: [0..3] queue heads (one per launch of a phased run), [4] error
// flag, then four per subband: F, CT, X and PCAL items done
#define FX_PCTR_ERR 4
#define FX_PCTR_FIRST 8
#define FX_PCTR_PER_SUB 4
#define FX_PNT 512         // threads per workgroup
#define FX_PXKT 32         // spectra per X stage (chan rows are padded to a multiple)
#define FX_PCT_T 128       // spectra per corner-turn item
// F groups of an nfft-point subband: 16384/nfft transforms per group
int fx_persist_group_items(int nfft);
// LDS bytes an item of this kind needs (F: nfft, with phase tables when
// pm; X: diagonal-only or any tile; PCAL: nfold; the others: none extra)
size_t fx_persist_lds(int type, int arg, bool pm);
// items [i0, i1) of the table on nblocks workgroups; wait: honour the items'
// waits and count their completions (one fused launch), else the caller
// orders the launches
int fxk_batch_persistent(hipStream_t s, const float2* tw, const fx_psub* subs, const fx_pitem* items, int i0, int i1,
                         int* ctrs, int head, int nblocks, int wait, size_t lds_bytes);
// workgroups that fill the device with this much dynamic LDS (0 on error)
int fxk_batch_persistent_blocks(size_t lds_bytes);

// the device fold and tone buffers and the extraction parameters of a
// phase-cal object, for kernels that fold and extract themselves (marks it
// finalized so fx_pcal_tones* read what they wrote); *position: its sample
// position modulo the fold length
int fx_pcal_device_state(fx_pcal* p, fx_psub* S, size_t* position);

// context internals shared with the task runtime and the application shims
const float2* fx_ctx_twiddles(fx_ctx* ctx);
hipStream_t fx_ctx_hip_stream(fx_ctx* ctx);
int fx_ctx_device_of(fx_ctx* ctx);

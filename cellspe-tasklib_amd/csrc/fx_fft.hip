// fx_fft.hip -- LDS-staged FFT kernels for gfx950 and the fused front ends.
//
//   k_fft_c2c        forward complex FFT (speFFT_exec / CELLDSP_TASK_FFT_EXEC,
//                    cellspe-tasklib-dsp.cpp:565-602)
//   k_fft_r2c        real-input FFT (genericFFTexec_R, architecture.h:563-578)
//   k_spectrometer   2-bit VDIF multi-channel unpack -> window -> real FFT ->
//                    |X|^2 accumulate (spectrometer_spu.c:97-224,253-349 with
//                    the IPP mean-power convention, TaskCoreIPP.cpp:437-511)
//   k_fstage         2-bit unpack -> real FFT -> FP64 phase model rotation ->
//                    channelized [c][t][i] store in FP32/FP16/FP8 (the F half
//                    of minicorrelator_spu.c:326-410 / exec_difxprocess,
//                    cellspe-tasklib-dsp.cpp:609-698)
//
// Every transform is a Stockham radix-16 FFT over a workgroup-resident LDS
// tile: F transforms of complex length M per workgroup with F*M = FM complex
// (8192 or 16384) and 512 threads, so each thread holds 16 (32) complex values
// per pass.  Real transforms of length N = 2M pack even/odd samples
// into one complex M-point FFT and split the result (rsplit).
#include "fx_internal.h"
#include "fx_fstage.h"

#include <hip/hip_fp8.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace {

template <int M>
struct FftShape {
    static constexpr int FM = (M >= 16384) ? 16384 : 8192;   // complex per workgroup tile
    static constexpr int F = FM / M;                          // transforms per tile
    static constexpr int NT = 512;                            // threads
    static constexpr int QPT = FM / NT;                       // complex per thread (16 or 32)
    static constexpr int SLT = FM / (2 * NT);                 // real-split pairs per thread
    static constexpr int LDS = lds_complex(FM);
};

// ---------------------------------------------------------------------------
// complex -> complex
// ---------------------------------------------------------------------------
// INV: the unnormalized backward transform (sign +) times scale, as
// conj(FFT(conj x)) -- the lag spectrum of a visibility row
// (zerocorr.c:391,633-645: FFTW_BACKWARD over the channels, 1/n_ffts).
template <int M, bool INV = false>
__global__ __launch_bounds__(FftShape<M>::NT) void k_fft_c2c(const float2* __restrict__ twg,
                                                             const float2* __restrict__ in,
                                                             float2* __restrict__ out, long batch,
                                                             float scale = 1.0f)
{
    using S = FftShape<M>;
    __shared__ float2 sh[S::LDS];
    __shared__ float2 tw[FX_TWL];
    const int tid = threadIdx.x;
    twl_load(tw, twg, tid, S::NT);
    const long f0 = (long)blockIdx.x * S::F;
    const long nvalid = (batch - f0 < S::F ? batch - f0 : S::F) * M;
#pragma unroll
    for (int s = 0; s < S::QPT; s++) {
        int e = tid + s * S::NT;
        float2 v = (e < nvalid) ? in[f0 * M + e] : make_float2(0.f, 0.f);
        if (INV) v.y = -v.y;
        sh[lpad(e)] = v;
    }
    __syncthreads();
    fft_passes<M, 1, S::NT, S::F>(sh, tw, tid);
#pragma unroll
    for (int s = 0; s < S::QPT; s++) {
        int e = tid + s * S::NT;
        if (e < nvalid) {
            float2 v = sh[lpad(e)];
            if (INV) v = make_float2(v.x * scale, -v.y * scale);
            out[f0 * M + e] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// real -> complex (N = 2M real samples per transform)
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(FftShape<M>::NT) void k_fft_r2c(const float2* __restrict__ twg,
                                                             const float2* __restrict__ in,
                                                             float2* __restrict__ out, long batch,
                                                             int keep_nyq)
{
    using S = FftShape<M>;
    __shared__ float2 sh[S::LDS];
    __shared__ float2 tw[FX_TWL];
    const int tid = threadIdx.x;
    twl_load(tw, twg, tid, S::NT);
    const long f0 = (long)blockIdx.x * S::F;
    const long nf = (batch - f0 < S::F ? batch - f0 : S::F);
    const long nvalid = nf * M;
#pragma unroll
    for (int s = 0; s < S::QPT; s++) {
        int e = tid + s * S::NT;
        sh[lpad(e)] = (e < nvalid) ? in[f0 * M + e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_passes<M, 1, S::NT, S::F>(sh, tw, tid);
    const int ostride = M + (keep_nyq ? 1 : 0);
    constexpr int H = (M / 2 > 0) ? M / 2 : 1;
    for (int q = tid; q < S::F * H; q += S::NT) {
        int f = q / H;
        int kp = q - f * H;
        if (f >= nf) continue;
        float2 xk, xm;
        rsplit<M>(sh, f * M, kp, tw, xk, xm);
        float2* o = out + (f0 + f) * ostride;
        o[kp] = xk;
        if (kp == 0) {
            if (keep_nyq) o[M] = xm;
        } else {
            o[M - kp] = xm;
        }
    }
    if (tid < S::F && tid < nf && M >= 2) {
        float2 xk, xm;
        rsplit<M>(sh, tid * M, M / 2, tw, xk, xm);
        out[(f0 + tid) * ostride + M / 2] = xk;
    }
}

// ---------------------------------------------------------------------------
// spectrometer: nchan-interleaved 2-bit words, one time sample per 2*nchan
// bits.  Workgroup (c, s, g) integrates FFTs [g*per, (g+1)*per) of spectrum s
// of channel c, F transforms at a time, and writes its partial power
// partial[wg][f][k].
// ---------------------------------------------------------------------------
// The spectrometer's tile: FftShape's, with 1024 threads for the 16384-point
// tile (one transform of M = 16384; its LDS allows one workgroup per CU, and
// 16 waves hide the load and LDS latency that 8 leave exposed).
template <int M>
struct SpecShape {
    static constexpr int FM = FftShape<M>::FM;
    static constexpr int F = FftShape<M>::F;
    static constexpr int NT = FM >= 16384 ? 1024 : FftShape<M>::NT;
    static constexpr int QPT = FM / NT;
    static constexpr int SLT = FM / (2 * NT);
    static constexpr int LDS = FftShape<M>::LDS;
};

template <int M>
__global__ __launch_bounds__(SpecShape<M>::NT) void k_spectrometer(
    const float2* __restrict__ twg, const float* __restrict__ win, const uint32_t* __restrict__ words,
    int nchan, int integrate_ffts, int groups, int per_group, long hop, long spec_stride,
    float* __restrict__ partial)
{
    using S = SpecShape<M>;
    __shared__ float2 sh[S::LDS];
    __shared__ float2 tw[FX_TWL];
    // the radix-16 passes' twiddle tables: W_256^{rk} (span 16) and
    // W_4096^{k}, W_4096^{4k}, W_4096^{8k} (span 256), as the F kernels use
    __shared__ float2 tw2[256];
    __shared__ float2 t3[768];
    const int tid = threadIdx.x;
    twl_load(tw, twg, tid, S::NT);
    for (int e = tid; e < 256; e += S::NT) tw2[e] = twg[(e >> 4) * (e & 15) * (FX_TW_N / 256)];
    for (int e = tid; e < 768; e += S::NT) {
        const int blk = e >> 8, k = e & 255;
        t3[e] = twg[(k << (blk == 0 ? 3 : blk == 1 ? 5 : 6)) & (FX_TW_N - 1)];
    }
    const int wg = blockIdx.x;
    const int g = wg % groups;
    const int cs = wg / groups;
    const int c = cs % nchan;
    const long s = cs / nchan;
    const int spw = 16 / nchan;          // time samples per word (a power of two)
    const int lspw = __builtin_ctz(spw);
    const int shift0 = 2 * c;
    const int bits_per_t = 2 * nchan;

    float acc[S::SLT][2];
#pragma unroll
    for (int i = 0; i < S::SLT; i++) acc[i][0] = acc[i][1] = 0.f;
    float accm = 0.f;

    int fft_count = per_group;
    if ((long)g * per_group + fft_count > integrate_ffts) fft_count = integrate_ffts - g * per_group;

    int tl = tid;
    for (int it = 0; it < fft_count; it += S::F) {
        // tl is tid re-derived every iteration (an empty asm the compiler
        // cannot see through): the split twiddles and addresses are
        // recomputed, not held across iterations in registers (or scratch)
        asm volatile("" : "+v"(tl));
        // load F transforms: F*N samples of channel c; thread handles 2*16 = 32
        // consecutive-in-pairs samples: sample pair e -> z[e] = (x[2e], x[2e+1])
        // (partial unroll bounds the loads in flight to the 1024-thread budget)
#pragma unroll 4
        for (int q = 0; q < S::QPT; q++) {
            int e = tid + q * S::NT;      // complex index within the WG tile
            int f = e / M;
            int n2 = 2 * (e - f * M);     // sample index within transform f
            float2 z = make_float2(0.f, 0.f);
            if (it + f < fft_count) {
                // FFT j of spectrum s starts at s*spec_stride + j*hop (hop = N/overlap)
                long t = s * spec_stride + ((long)g * per_group + it + f) * hop + n2;
                uint32_t w0, w1;
                if (spw == 1 && ((t & 1) == 0) && (((uintptr_t)words & 7) == 0)) {
                    // 16 channels: one word a time sample, both words at once
                    const uint2 ww = *reinterpret_cast<const uint2*>(words + t);
                    w0 = ww.x;
                    w1 = ww.y;
                } else {
                    w0 = words[t >> lspw];
                    w1 = words[(t + 1) >> lspw];
                }
                int p0 = (int)(t & (spw - 1)) * bits_per_t + shift0;
                int p1 = (int)((t + 1) & (spw - 1)) * bits_per_t + shift0;
                z.x = lut2((w0 >> p0) & 3u) * win[n2];
                z.y = lut2((w1 >> p1) & 3u) * win[n2 + 1];
            }
            sh[lpad(e)] = z;
        }
        __syncthreads();
        fft_passes_hook<M, 1, S::NT, S::F, true, true>(sh, tw, tid, tw2, t3, [](auto) {});
        // power of the thread's 8 (kp, f) items, kp-major so the same thread
        // owns the same bins on every iteration
#pragma unroll
        for (int i = 0; i < S::SLT; i++) {
            int q = tl * S::SLT + i;
            int kp = q / S::F;
            int f = q - kp * S::F;
            if (it + f < fft_count) {
                float2 xk, xm;
                rsplit<M>(sh, f * M, kp, tw, xk, xm);
                acc[i][0] += xk.x * xk.x + xk.y * xk.y;
                acc[i][1] += xm.x * xm.x + xm.y * xm.y;
            }
        }
        if (tid < S::F && it + tid < fft_count) {
            float2 xk, xm;
            rsplit<M>(sh, tid * M, M / 2, tw, xk, xm);
            accm += xk.x * xk.x + xk.y * xk.y;
        }
        __syncthreads();
    }
    // partial rows hold M+1 bins: k = M is the Nyquist bin (the kp = 0 pair's X_{M})
    float* pw = partial + (long)wg * S::F * (M + 1);
#pragma unroll
    for (int i = 0; i < S::SLT; i++) {
        int q = tid * S::SLT + i;
        int kp = q / S::F;
        int f = q - kp * S::F;
        pw[f * (M + 1) + kp] = acc[i][0];
        pw[f * (M + 1) + (M - kp)] = acc[i][1];
    }
    if (tid < S::F) pw[tid * (M + 1) + M / 2] = accm;
}

// out[s][c][k] = (1/K) sum_g sum_f partial[(s*nchan+c)*groups+g][f][k]
__global__ void k_spec_reduce(const float* __restrict__ partial, int M, int F, int groups,
                              int nchan, long nspec, float scale, int nyq, float* __restrict__ out)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int MO = M + nyq;   // output bins per spectrum
    long total = nspec * nchan * (long)MO;
    if (idx >= total) return;
    int k = (int)(idx % MO);
    long cs = idx / MO;
    const float* p = partial + cs * (long)groups * F * (M + 1);
    float sum = 0.f;
    for (int g = 0; g < groups; g++)
        for (int f = 0; f < F; f++) sum += p[((long)g * F + f) * (M + 1) + k];
    out[idx] = sum * scale;
}

// sc_chan: per-channel scales, or nullptr when every channel has the scale
// sc_uni; sc_inp: per-input scales or nullptr.  PMO: launched with a phase
// model, no multiplier and no per-channel scales.
template <int M, typename T, bool PMO>
__global__ __launch_bounds__(FftShape<M>::NT) void k_fstage(
    const float2* __restrict__ twg, const uint32_t* __restrict__ words, size_t in_stride, int ninp,
    long t_first, long nt, const fx_phase_model* __restrict__ pm, const float2* __restrict__ mult,
    const float* __restrict__ sc_chan, float sc_uni, const float* __restrict__ sc_inp, T* __restrict__ nat)
{
    using S = FftShape<M>;
    constexpr int F = S::F;
    constexpr int NT = S::NT;
    constexpr int N = 2 * M;
    typedef typename ChanT<T>::V V;
    // item order: g = t*I + i (FP32/FP16 engines, [c][t][i] after the corner
    // turn) or g = i*nt + t (FP8 engine, [c][i][t])
    constexpr bool IM = std::is_same<T, uint16_t>::value;
    constexpr int PLO = PhaseTab<M>::PLO;
    __shared__ float2 sh[S::LDS];
    __shared__ FTabs<M> tb;
    __shared__ FItems<M, F> it;
    const int tid = threadIdx.x;
    const long g0 = (long)blockIdx.x * F;
    fstage_front<M, F, NT, IM>(twg, words, in_stride, ninp, t_first, nt, pm, sc_uni, sc_inp, g0, sh, tb, it, tid);
    const bool has_pm = PMO || pm != nullptr;

    // ---- split, rotate, scale, convert; lanes write consecutive channels of
    // one row
    V* o = reinterpret_cast<V*>(nat);
    constexpr int H = M / 2;
    if constexpr (H >= NT && NT % PLO == 0) {
        // thread tid owns channels kp = q NT + tid of every item: the split
        // twiddle W_N^kp advances by W_N^NT per q (shared by the items), and
        // the low phasor factors of kp and M - kp are the same for every q
        const float2 wstep = twg[NT * (FX_TW_N / N)];
        const float2 w0 = twg[tid * (FX_TW_N / N)];
        const int lo_k = tid % PLO, lo_m = (PLO - lo_k) % PLO;
        // blocks whose items all exist (all but the last) skip the per-item test
        auto body = [&](auto checked) {
        float2 w = w0;
#pragma unroll
        for (int q = 0; q < H / NT; q++) {
            const int kp = q * NT + tid;
            if (q) w = cmul(w, wstep);
            const int k2 = M - kp;
            const int k2m = k2 & (M - 1);   // kp = 0: k2 = M is the dropped Nyquist bin
#pragma unroll
            for (int f = 0; f < F; f++) {
                if constexpr (decltype(checked)::value) {
                    if (it.t[f] < 0) continue;
                }
                V* row = o + (size_t)(g0 + f) * M;
                float2 xk, xm;
                rsplit2_tw<M>(sh, f * M, kp, w, xk, xm);
                row[kp] = fstage_finish<M, T, F, PMO>(xk, it, f, kp, has_pm ? it.ph_lo[f][lo_k] : make_float2(1.f, 0.f),
                                                 has_pm, mult, sc_chan);
                if (kp != 0)
                    row[k2] = fstage_finish<M, T, F, PMO>(xm, it, f, k2m,
                                                     has_pm ? it.ph_lo[f][lo_m] : make_float2(1.f, 0.f), has_pm,
                                                     mult, sc_chan);
            }
        }
        };
        if (it.t[F - 1] >= 0)
            body(std::false_type{});
        else
            body(std::true_type{});
    } else {
#pragma unroll
        for (int q = 0; q < (F * H) / NT; q++) {
            const int f = H >= NT ? (q * NT) / H : (q * NT + tid) / H;
            const int kp = H >= NT ? (q * NT) % H + tid : (q * NT + tid) % H;
            if (it.t[f] < 0) continue;
            float2 xk, xm;
            rsplit2_tw<M>(sh, f * M, kp, twg[kp * (FX_TW_N / N)], xk, xm);
            const int k2 = M - kp;
            const int k2m = k2 & (M - 1);
            V* row = o + (size_t)(g0 + f) * M;
            row[kp] = fstage_finish<M, T, F, PMO>(xk, it, f, kp, has_pm ? it.ph_lo[f][kp % PLO] : make_float2(1.f, 0.f),
                                             has_pm, mult, sc_chan);
            if (kp != 0)
                row[k2] = fstage_finish<M, T, F, PMO>(xm, it, f, k2m,
                                                 has_pm ? it.ph_lo[f][k2m % PLO] : make_float2(1.f, 0.f), has_pm,
                                                 mult, sc_chan);
        }
    }
    // kp = M/2 of each transform
    if (tid < F && it.t[tid] >= 0) {
        float2 xk, xm;
        rsplit2_tw<M>(sh, tid * M, M / 2, twg[(M / 2) * (FX_TW_N / N)], xk, xm);
        o[(size_t)(g0 + tid) * M + M / 2] = fstage_finish<M, T, F, PMO>(
            xk, it, tid, M / 2, has_pm ? it.ph_lo[tid][(M / 2) % PLO] : make_float2(1.f, 0.f), has_pm, mult, sc_chan);
    }
}

// Store of F consecutive channelized values as 16-byte (or 8-byte) vectors.
template <typename V, int F>
__device__ __forceinline__ void store_run(V* dst, const V (&v)[F])
{
    constexpr int B = (int)sizeof(V) * F;
    if constexpr (B % 16 == 0) {
#pragma unroll
        for (int j = 0; j < B / 16; j++)
            reinterpret_cast<uint4*>(dst)[j] = reinterpret_cast<const uint4*>(v)[j];
    } else if constexpr (B == 8) {
        *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(v);
    } else {
#pragma unroll
        for (int f = 0; f < F; f++) dst[f] = v[f];
    }
}

// F kernel with the corner turn folded in: a 1024-thread workgroup transforms
// F = 16384/M items per round, RD rounds, FC = RD*F consecutive columns of
// the engine layout (FC inputs of one spectrum for [c][t][i], or FC spectra
// of one input for the FP8 [c][i][t]).  The values of every round stay in
// registers and each lane stores one channel of all FC columns as one
// vector out[k][col .. col+FC-1] (RD = 2 for the 4-byte FP16 and 2-byte FP8
// samples: 32- and 16-byte runs, so the vector memory unit touches half as
// many lines).  The workgroups of neighbouring columns run on one XCD
// (xcd_block), so the partial lines of a spectrum meet in that XCD's L2 and
// reach HBM as (mostly) whole lines.  The host guarantees that every
// workgroup's FC items exist and that col0, ostride and ostride_i are
// multiples of FC.
template <int M, typename T, bool PMO, int RD, int FMX>
__global__ __launch_bounds__(FMX / 16) void k_fstage_direct(
    const float2* __restrict__ twg, const uint32_t* __restrict__ words, size_t in_stride, int ninp,
    long t_first, long nt, const fx_phase_model* __restrict__ pm, const float2* __restrict__ mult,
    const float* __restrict__ sc_chan, float sc_uni, const float* __restrict__ sc_inp,
    typename ChanT<T>::V* __restrict__ out, long ostride, long col0, long ostride_i)
{
    constexpr int NT = FMX / 16;
    constexpr int F = FMX / M;
    constexpr int FC = F * RD;
    constexpr int N = 2 * M;
    constexpr int H = M / 2;
    constexpr int QN = H / NT;
    static_assert(H % NT == 0 && NT % 64 == 0, "direct F kernel: M in 2048 .. 8192");
    typedef typename ChanT<T>::V V;
    constexpr bool IM = std::is_same<T, uint16_t>::value;
    constexpr int PLO = PhaseTab<M>::PLO;
    __shared__ float2 sh[lds_complex(F * M)];
    __shared__ FTabs<M> tb;
    __shared__ FItems<M, F> it;
    const int tid = threadIdx.x;
    const long nblk = nt * ninp / FC;
    const long lb = xcd_block(nblk);
    if (lb >= nblk) return;
    const long gb = lb * FC;
    const bool has_pm = PMO || pm != nullptr;
    const long col = IM ? (gb / nt) * ostride_i + col0 + (gb % nt) : col0 + gb;
    const int lo_k = tid % PLO, lo_m = (PLO - lo_k) % PLO;
    V vk[QN][FC], vm[QN][FC];
    // tl is tid re-derived every round (an empty asm the compiler cannot see
    // through), so the rounds' address arithmetic is not shared and kept live
    // beside the held values
    int tl = tid;
#pragma unroll
    for (int rd = 0; rd < RD; rd++) {
        if (rd) {
            __syncthreads();   // the previous round's epilogue has read sh and it
            asm volatile("" : "+v"(tl));
        }
        // the first round also fills the LDS tables
        fstage_group<M, F, NT, IM>(rd == 0 ? twg : nullptr, words, in_stride, ninp, t_first, nt, pm, sc_uni, sc_inp,
                                   gb + rd * F, sh, tb, it, tl, [](auto) {});
        const float2 wstep = twg[NT * (FX_TW_N / N)];
        float2 w = twg[tl * (FX_TW_N / N)];
#pragma unroll
        for (int q = 0; q < QN; q++) {
            const int kp = q * NT + tl;
            if (q) w = cmul(w, wstep);
            const int k2m = (M - kp) & (M - 1);   // kp = 0: the dropped Nyquist bin
#pragma unroll
            for (int f = 0; f < F; f++) {
                float2 xk, xm;
                rsplit2_tw<M>(sh, f * M, kp, w, xk, xm);
                vk[q][rd * F + f] = fstage_finish<M, T, F, PMO>(
                    xk, it, f, kp, has_pm ? it.ph_lo[f][lo_k] : make_float2(1.f, 0.f), has_pm, mult, sc_chan);
                vm[q][rd * F + f] = fstage_finish<M, T, F, PMO>(
                    xm, it, f, k2m, has_pm ? it.ph_lo[f][lo_m] : make_float2(1.f, 0.f), has_pm, mult, sc_chan);
            }
        }
        // kp = M/2 of each item
        if (tl < F) {
            float2 xk, xm;
            rsplit2_tw<M>(sh, tl * M, M / 2, twg[(M / 2) * (FX_TW_N / N)], xk, xm);
            out[(size_t)(M / 2) * ostride + col + rd * F + tl] = fstage_finish<M, T, F, PMO>(
                xk, it, tl, M / 2, has_pm ? it.ph_lo[tl][(M / 2) % PLO] : make_float2(1.f, 0.f), has_pm, mult,
                sc_chan);
        }
    }
#pragma unroll
    for (int q = 0; q < QN; q++) {
        const int kp = q * NT + tid;
        store_run<V, FC>(out + (size_t)kp * ostride + col, vk[q]);
        if (kp != 0) store_run<V, FC>(out + (size_t)(M - kp) * ostride + col, vm[q]);
    }
}

// The value of the other lane of each lane pair (2m, 2m+1): a DPP quad
// permutation [1, 0, 3, 2].  One-word channel values only (FP16 complex);
// the generic form is never called.
template <typename V>
__device__ __forceinline__ V swap_lane_pairs(V v)
{
    return v;
}
template <>
__device__ __forceinline__ uint32_t swap_lane_pairs<uint32_t>(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

// Inputs of the persistent F kernel: their phase models and scales are held
// in LDS (more inputs take the one-block-per-workgroup kernel).
#define FX_PDIR_PM 256

// Persistent form of k_fstage_direct (one round: F = FMX/M items, the F
// inputs of one spectrum, per block).  Workgroup j of XCD x (blockIdx = 8 j
// + x, one 1024-thread workgroup per CU) takes blocks lo + j, lo + j + W,
// lo + j + 2W, ... of the XCD's contiguous range [lo, hi) (W workgroups per
// XCD), so the XCD's workgroups advance through its range side by side and
// the blocks whose 16-byte runs share the 128-byte lines of a spectrum run
// together on one XCD (their partial lines merge in its L2), as with one block
// per workgroup.  What it removes: a workgroup's epilogue ends in line stores
// that take longer to issue than to compute, and with one block per
// workgroup the CU idles until they have drained and the next workgroup has
// started (the LDS holds one workgroup per CU).  Here the next block's packed
// words are read during the current block's epilogue, and that epilogue's
// stores drain while the next block transforms: inside the loop every
// barrier waits for LDS operations only, and nothing else waits for a global
// access (phase models and input scales come from LDS, the epilogue's
// twiddles from registers; the first block is peeled off the loop so the wait
// for the prefetched words never covers the previous block's stores).
// HASPM: a phase model is applied; PMO: phase model only (no multiplier, no
// per-channel scales).
template <int M, typename T, bool HASPM, bool PMO, bool PAIR, int FMX>
__global__ __launch_bounds__(FMX / 16) void k_fstage_direct_p(
    const float2* __restrict__ twg, const uint32_t* __restrict__ words, size_t in_stride, int ninp,
    long t_first, long nt, const fx_phase_model* __restrict__ pm, const float2* __restrict__ mult,
    const float* __restrict__ sc_chan, float sc_uni, const float* __restrict__ sc_inp,
    typename ChanT<T>::V* __restrict__ out, long ostride, long col0, unsigned* __restrict__ ctr)
{
    constexpr int NT = FMX / 16;
    constexpr int F = FMX / M;
    constexpr int N = 2 * M;
    constexpr int H = M / 2;
    constexpr int QN = H / NT;
    constexpr int WPT = M / 8;               // packed words per transform
    constexpr int WTH = (F * M / 8) / NT;    // packed words per thread
    constexpr int PLO = PhaseTab<M>::PLO;
    static_assert(H % NT == 0 && NT % 64 == 0 && WTH % 2 == 0, "persistent direct F kernel: M in 2048 .. 8192");
    static_assert(!PMO || HASPM, "phase-model-only needs a phase model");
    static_assert(!PAIR || std::is_same<T, __half2>::value, "paired stores: FP16 channels (one word a sample)");
    constexpr int U = PAIR ? 2 : 1;          // blocks per unit of the queue
    typedef typename ChanT<T>::V V;
    __shared__ float2 sh[lds_complex(F * M)];
    __shared__ FTabs<M> tb;
    __shared__ FItems<M, F> it;
    __shared__ fx_phase_model pms[HASPM ? FX_PDIR_PM : 1];
    __shared__ float sis[FX_PDIR_PM];
    __shared__ long nextb;
    const int tid = threadIdx.x;
    const int x = (int)(blockIdx.x & 7);
    const long nblk = nt * ninp / (F * U);   // units
    const long per = (nblk + 7) / 8;
    const long lo = x * per;
    const long hi = lo + per < nblk ? lo + per : nblk;
    for (int i = tid; i < ninp; i += NT) {
        if constexpr (HASPM) pms[i] = pm[i];
        sis[i] = sc_inp ? sc_inp[i] : 1.0f;
    }
    if (tid == 0) nextb = lo + (long)atomicAdd(&ctr[x], 1u);
    lds_barrier();   // every thread reads other inputs' models and scales, and the first block
    long b = nextb;   // the unit: blocks U b .. U b + U - 1
    if (b >= hi) return;
    // this thread's staged words: item wf of a block, words wof .. wof+WTH-1
    const int wl = WTH * tid;
    const int wf = wl / WPT;
    const int wof = wl - wf * WPT;
    uint2 wpre[WTH / 2];
    // (blocks are the same in every thread: the block's spectrum and first
    // input come from one scalar division; its items are inputs i0 .. i0+F-1)
    auto wload = [&](long blk) {
        const long g = uniform_long(blk) * F;
        const long t = g / ninp;
        const int i = (int)(g - t * ninp) + wf;
        const uint32_t* src = words + (size_t)i * in_stride + (size_t)(t_first + t) * WPT + wof;
#pragma unroll
        for (int q = 0; q < WTH; q += 2) wpre[q / 2] = *reinterpret_cast<const uint2*>(src + q);
    };
    wload(U * b);
    FConstsPre<M, NT> cpre;
    cpre.load(twg, tid);
    // the epilogue's split twiddles W_N^kp, kp = q NT + tid: the first, the
    // step per q and (threads < F) that of kp = M/2; held in registers, or
    // (PAIR, where the held channel values need the registers) read from
    // the table in each epilogue, before its word reads so the wait for them
    // leaves those in flight
    float2 w0r, wstepr, wmidr;
    auto twiddles = [&]() {
        w0r = twg[tid * (FX_TW_N / N)];
        wstepr = twg[NT * (FX_TW_N / N)];
        wmidr = twg[(M / 2) * (FX_TW_N / N)];
    };
    if constexpr (!PAIR) twiddles();
    const fx_phase_model* pml = HASPM ? pms : nullptr;
    int tl = tid;
    // thread 0: the next block's offset in the XCD's range.  Left undefined
    // in the other threads (they never read it), so no value merges with the
    // atomic's result where its branch ends and the wait for the result sits
    // at its one use, two passes later.
    unsigned nb;
    // PAIR: the first block's channel values, held until the second block's
    // (one word each: [q][channel k, M - k][item])
    V hold[PAIR ? QN : 1][2][F];
    // sub = 0: a lone block, stored at once; 1: a pair's first block, held;
    // 2: a pair's second block, stored with the first as 32-byte runs
    // (both lambdas inlined: an out-of-line call would put the arrays they
    // capture by reference in scratch memory)
    auto block = [&](auto first, auto sub) __attribute__((always_inline)) -> long {
        constexpr int SUB = decltype(sub)::value;
        const long gb = (U * b + (SUB == 2 ? 1 : 0)) * F;
        // an increment (wrapping at 2^32 - 1), which the compiler issues as
        // one atomic and does not rewrite into a wave-combined add whose
        // result is needed at once
        if (SUB != 2 && tid == 0) nb = atomicInc(&ctr[x], 0xffffffffu);
        fstage_group<M, F, NT, false, true>(words, in_stride, ninp, t_first, nt, pml, sc_uni, sis, gb, sh, tb, it, tl,
                                      [&](auto p) {
            // published after the second pass, visible after the third's barrier
            if constexpr (decltype(p)::value == 1 && SUB != 2) {
                if (tid == 0) nextb = lo + (long)nb;
            }
        },
                                      []() {},
                                      [&](uint32_t* raw) {
#pragma unroll
            for (int q = 0; q < WTH; q += 2) *reinterpret_cast<uint2*>(raw + wl + q) = wpre[q / 2];
            if constexpr (decltype(first)::value) cpre.put(tb, tl);
        });
        // the next block's words (past the range: this block's again, unused)
        // arrive during this epilogue
        const long bn = nextb;
        if constexpr (PAIR) twiddles();
        if constexpr (SUB == 1)
            wload(U * b + 1);
        else
            wload(U * (bn < hi ? bn : b));
        const float2 w0 = w0r, wstep = wstepr, wmid = wmidr;
        const long col = col0 + gb;
        // kp = M/2 of each item (first: the stores every lane issues come
        // last, so the next block's wait for its words leaves them in flight)
        if (tl < F) {
            float2 xk, xm;
            rsplit2_tw<M>(sh, tl * M, M / 2, wmid, xk, xm);
            out[(size_t)(M / 2) * ostride + col + tl] = fstage_finish<M, T, F, PMO>(
                xk, it, tl, M / 2, HASPM ? it.ph_lo[tl][(M / 2) % PLO] : make_float2(1.f, 0.f), HASPM, mult, sc_chan);
        }
        float2 w = w0;
#pragma unroll
        for (int q = 0; q < QN; q++) {
            // PAIR: the phasor factors are read again for every q (tq hides
            // that they are the same ones), so they are not held across q
            // beside the held channel values
            int tq = tl;
            if constexpr (PAIR) asm volatile("" : "+v"(tq));
            const int lo_k = tq % PLO, lo_m = (PLO - lo_k) % PLO;
            const int kp = q * NT + tl;
            if (q) w = cmul(w, wstep);
            const int k2m = (M - kp) & (M - 1);   // kp = 0: the dropped Nyquist bin
            V vk[F], vm[F];
#pragma unroll
            for (int f = 0; f < F; f++) {
                float2 xk, xm;
                rsplit2_tw<M>(sh, f * M, kp, w, xk, xm);
                vk[f] = fstage_finish<M, T, F, PMO>(xk, it, f, kp, HASPM ? it.ph_lo[f][lo_k] : make_float2(1.f, 0.f),
                                                   HASPM, mult, sc_chan);
                vm[f] = fstage_finish<M, T, F, PMO>(xm, it, f, k2m, HASPM ? it.ph_lo[f][lo_m] : make_float2(1.f, 0.f),
                                                   HASPM, mult, sc_chan);
            }
            // kp = 0 has no channel M - kp: that lane stores its channel 0 a
            // second time (same address, same values), so every lane issues
            // the store
#pragma unroll
            for (int f = 0; f < F; f++)
                if (kp == 0) vm[f] = vk[f];
            if constexpr (SUB == 0) {
                store_run<V, F>(out + (size_t)kp * ostride + col, vk);
                store_run<V, F>(out + (size_t)(kp ? M - kp : 0) * ostride + col, vm);
            } else if constexpr (SUB == 1) {
#pragma unroll
                for (int f = 0; f < F; f++) {
                    hold[q][0][f] = vk[f];
                    hold[q][1][f] = vm[f];
                }
            } else {
                // Lanes 2m and 2m+1 hold channels kp and kp ^ 1 of both
                // blocks (the first block's F columns, then the second's):
                // they swap the first block's value of the odd channel for
                // the second block's value of the even one, then each store
                // writes one channel's 2F columns as two neighbouring 16-byte
                // halves, one per lane (one 32-byte piece of the line).
                // Row offsets: the even channel kpe = q NT + (tl & ~1) from
                // one per-block product and a uniform step per q; the odd
                // one a row further (channel kpe + 1) or nearer (M - kpe - 1);
                // kpe = 0 has no mirror: its lanes store channel 0 again.
                const bool odd = (tl & 1) != 0;
                const size_t os = (size_t)ostride;
                const size_t ke = (size_t)q * NT * os + (size_t)(tl & ~1) * os;
                V* const pb = out + (col - F) + (odd ? F : 0);   // the pair's first column, + F for the odd lane
                V* const pke = pb + ke;
                V* const pmo = pb + (size_t)M * os - ke - os;
                V* const pme = (q == 0 && tl < 2) ? pb : pb + (size_t)M * os - ke;
#pragma unroll
                for (int s2 = 0; s2 < 2; s2++) {
                    V st1[F], st2[F];
#pragma unroll
                    for (int f = 0; f < F; f++) {
                        const V a = hold[q][s2][f];
                        const V bv = s2 ? vm[f] : vk[f];
                        const V recv = swap_lane_pairs(odd ? a : bv);
                        st1[f] = odd ? recv : a;
                        st2[f] = odd ? bv : recv;
                    }
                    store_run<V, F>(s2 ? pme : pke, st1);
                    store_run<V, F>(s2 ? pmo : pke + os, st2);
                }
            }
        }
        return bn;
    };
    auto unit = [&](auto first) __attribute__((always_inline)) -> long {
        if constexpr (PAIR) {
            block(first, std::integral_constant<int, 1>{});
            lds_barrier();   // the first block's epilogue has read sh and it
            asm volatile("" : "+v"(tl));
            return block(std::false_type{}, std::integral_constant<int, 2>{});
        } else {
            return block(first, std::integral_constant<int, 0>{});
        }
    };
    b = unit(std::true_type{});
    while (b < hi) {
        lds_barrier();   // this epilogue's reads of sh and it precede the next block's staging
        // tl is tid re-derived every block (an empty asm the compiler cannot
        // see through): the block's LDS address arithmetic is not hoisted into
        // registers held across blocks
        asm volatile("" : "+v"(tl));
        b = unit(std::false_type{});
    }
}

// Corner turn of one sub-chunk of F-stage rows nat[g][c] (g < nitems) into
// the chunk-wide engine layout out[c][col], through a TL x TL LDS tile: both
// the 256-thread reads (rows of nat) and the writes (rows of out) are
// contiguous runs of TL elements.  Row c of out holds ostride columns.
// t-major items (g = t'*I + i) land at col = col0 + g; input-major items
// (g = i*ns + t') at col = i*ostride_i + col0 + t' (ostride_i = spectra per
// chunk).  XCD-aware: tiles of neighbouring items share an XCD.  (Used when
// the direct F kernel does not apply: other FFT lengths or input counts.)
template <typename V, int TL>
__global__ __launch_bounds__(256) void k_cornerturn(const V* __restrict__ nat, V* __restrict__ out,
                                                    long nitems, int C, long ostride, long col0,
                                                    int imajor, int ns, long ostride_i)
{
    __shared__ V tile[TL][TL + 1];
    constexpr int RS = 256 / TL;   // rows per thread step
    const long ntg = (nitems + TL - 1) / TL;
    const long ntc = C / TL;
    const long lb = xcd_block(ntg * ntc);
    if (lb >= ntg * ntc) return;
    const long tg = lb % ntg;
    const int tc = (int)(lb / ntg);
    const int lx = threadIdx.x % TL, ly = threadIdx.x / TL;
    const long gbase = tg * TL;
    const int cbase = tc * TL;
#pragma unroll
    for (int r = ly; r < TL; r += RS) {
        long g = gbase + r;
        if (g < nitems) tile[r][lx] = nat[(size_t)g * C + cbase + lx];
    }
    __syncthreads();
    const long g = gbase + lx;
    const long col = imajor ? (g / ns) * ostride_i + col0 + (g % ns) : col0 + g;
#pragma unroll
    for (int r = ly; r < TL; r += RS) {
        if (g < nitems) out[(size_t)(cbase + r) * ostride + col] = tile[lx][r];
    }
}

template <int M>
int launch_c2c(hipStream_t s, const float2* tw, const float2* in, float2* out, long batch, int inverse = 0,
               float scale = 1.0f)
{
    using S = FftShape<M>;
    long nblk = (batch + S::F - 1) / S::F;
    if (inverse)
        hipLaunchKernelGGL((k_fft_c2c<M, true>), dim3((unsigned)nblk), dim3(S::NT), 0, s, tw, in, out, batch,
                           scale);
    else
        hipLaunchKernelGGL((k_fft_c2c<M, false>), dim3((unsigned)nblk), dim3(S::NT), 0, s, tw, in, out, batch,
                           1.0f);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

template <int M>
int launch_r2c(hipStream_t s, const float2* tw, const float* in, float2* out, long batch, int keep)
{
    using S = FftShape<M>;
    long nblk = (batch + S::F - 1) / S::F;
    hipLaunchKernelGGL(k_fft_r2c<M>, dim3((unsigned)nblk), dim3(S::NT), 0, s, tw,
                       reinterpret_cast<const float2*>(in), out, batch, keep);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

template <int M>
int launch_spec(hipStream_t s, const float2* tw, const float* win, const uint32_t* words, long hop,
                long spec_stride, int nchan, int integrate_ffts, int nspec, int groups, int nyq, float* partial,
                float* out)
{
    using S = SpecShape<M>;
    int per_group = (integrate_ffts + groups - 1) / groups;
    per_group = ((per_group + S::F - 1) / S::F) * S::F;
    long nblk = (long)nspec * nchan * groups;
    hipLaunchKernelGGL(k_spectrometer<M>, dim3((unsigned)nblk), dim3(S::NT), 0, s, tw, win, words,
                       nchan, integrate_ffts, groups, per_group, hop, spec_stride, partial);
    if (hipGetLastError() != hipSuccess) return FX_ERR_HIP;
    long total = (long)nspec * nchan * (M + nyq);
    hipLaunchKernelGGL(k_spec_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       partial, M, S::F, groups, nchan, (long)nspec, 1.0f / (float)integrate_ffts, nyq,
                       out);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

template <typename V>
int launch_cornerturn_t(hipStream_t s, const void* nat, void* out, long nitems, int M, long ostride, long col0,
                        int imajor, int ns, long ostride_i)
{
    long ntiles = ((nitems + 63) / 64) * (M / 64);
    hipLaunchKernelGGL((k_cornerturn<V, 64>), dim3((unsigned)xcd_grid(ntiles)), dim3(256), 0, s,
                       reinterpret_cast<const V*>(nat), reinterpret_cast<V*>(out), nitems, M, ostride, col0,
                       imajor, ns, ostride_i);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

// out == nullptr: the F kernel alone (natural rows), no corner turn
template <int M, typename T>
int launch_fstage_t(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride,
                    int ninp, long t_first, long nt, const fx_phase_model* pm, const float2* mult,
                    const float* sc, float su, const float* si, void* nat, void* out)
{
    using S = FftShape<M>;
    typedef typename ChanT<T>::V V;
    long nitems = nt * ninp;
    long nblk = (nitems + S::F - 1) / S::F;
    if (pm && !mult && !sc)
        hipLaunchKernelGGL((k_fstage<M, T, true>), dim3((unsigned)nblk), dim3(S::NT), 0, s, tw, words, stride,
                           ninp, t_first, nt, pm, mult, sc, su, si, reinterpret_cast<T*>(nat));
    else
        hipLaunchKernelGGL((k_fstage<M, T, false>), dim3((unsigned)nblk), dim3(S::NT), 0, s, tw, words, stride,
                           ninp, t_first, nt, pm, mult, sc, su, si, reinterpret_cast<T*>(nat));
    if (hipGetLastError() != hipSuccess) return FX_ERR_HIP;
    if (!out) return FX_OK;
    return launch_cornerturn_t<V>(s, nat, out, nitems, M, nitems, 0, std::is_same<T, uint16_t>::value ? 1 : 0,
                                  (int)nt, nt);
}

template <int M>
int launch_fstage(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp,
                  long t_first, long nt, const fx_phase_model* pm, const float2* mult,
                  const float* sc, float su, const float* si, int out_type, void* nat, void* out)
{
    switch (out_type) {
    case FX_PREC_F32:
        return launch_fstage_t<M, float2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, nat, out);
    case FX_PREC_F16:
        return launch_fstage_t<M, __half2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, nat, out);
    case FX_PREC_FP8:
        return launch_fstage_t<M, uint16_t>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, nat, out);
    default:
        return FX_ERR_ARG;
    }
}

// rounds of the direct F kernel: 2 for the 2-byte FP8 samples (16-byte runs
// per lane: 7.10 vs 7.31 ms per 4096-spectrum chunk of 128 inputs), 1 for
// FP16 and FP32 (16- and 32-byte runs; two FP16 rounds measured 9.30 vs
// 8.63 ms, their held values leave the straight-line epilogue no registers)
template <typename T>
constexpr int direct_rounds()
{
    return std::is_same<T, uint16_t>::value ? 2 : 1;
}

// complex points per workgroup round of the direct F kernel (threads = /16):
// 16384 (1024 threads, one workgroup per CU).  The store width per lane is
// what matters: 8192-point rounds (512 threads, two workgroups per CU, 8-byte
// FP16 runs) measured 13.95 vs 8.63 ms per chunk.
template <typename T>
constexpr int direct_fm()
{
    return 16384;
}

template <int M, typename T>
int launch_fstage_direct_t(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp,
                           long t_first, long nt, const fx_phase_model* pm, const float2* mult, const float* sc, float su,
                           const float* si, void* out, long ostride, long col0, long ostride_i)
{
    constexpr int RD = direct_rounds<T>();
    constexpr int FMX = direct_fm<T>();
    constexpr int FC = (FMX / M) * RD;
    constexpr bool IM = std::is_same<T, uint16_t>::value;
    if ((IM ? nt : (long)ninp) % FC || ostride % FC || col0 % FC || (IM && ostride_i % FC)) return FX_ERR_ARG;
    const long nblk = nt * ninp / FC;
    if (nblk == 0) return FX_OK;
    const long grid = xcd_grid(nblk);
    // the straight-line phase-model-only epilogue needs more registers than
    // two rounds leave (it spills there), so two-round launches take the
    // general one
    if (RD == 1 && pm && !mult && !sc)
        hipLaunchKernelGGL((k_fstage_direct<M, T, RD == 1, RD, FMX>), dim3((unsigned)grid), dim3(FMX / 16), 0, s, tw,
                           words, stride, ninp, t_first, nt, pm, mult, sc, su, si,
                           reinterpret_cast<typename ChanT<T>::V*>(out), ostride, col0, ostride_i);
    else
        hipLaunchKernelGGL((k_fstage_direct<M, T, false, RD, FMX>), dim3((unsigned)grid), dim3(FMX / 16), 0, s, tw,
                           words, stride, ninp, t_first, nt, pm, mult, sc, su, si,
                           reinterpret_cast<typename ChanT<T>::V*>(out), ostride, col0, ostride_i);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

// the persistent form: nwg = workgroups to launch (one per CU; at least 8);
// pair: FP16 blocks in pairs where the inputs allow (32-byte store runs)
template <int M, typename T>
int launch_fstage_direct_p(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp,
                           long t_first, long nt, const fx_phase_model* pm, const float2* mult, const float* sc,
                           float su, const float* si, void* out, long ostride, long col0, unsigned* ctr, int nwg,
                           bool pair)
{
    constexpr int FMX = direct_fm<T>();
    constexpr int F = FMX / M;
    if (ninp % F || ostride % F || col0 % F || nwg < 8 || ninp > FX_PDIR_PM) return FX_ERR_ARG;
    const long nblk = nt * ninp / F;
    if (nblk == 0) return FX_OK;
    const long grid = std::min((long)(nwg / 8) * 8, xcd_grid(nblk));
    if (fxk_fill32(s, ctr, 0u, 8) != FX_OK) return FX_ERR_HIP;
    auto o = reinterpret_cast<typename ChanT<T>::V*>(out);
    // FP16 with the blocks of a spectrum in pairs: 32-byte store runs
    constexpr bool P2 = std::is_same<T, __half2>::value;
    if (P2 && pair && ninp % (2 * F) == 0) {
        // the general epilogue also for a phase model alone: the straight-line
        // one needs more registers than the held values leave (it spills)
        if (pm)
            hipLaunchKernelGGL((k_fstage_direct_p<M, T, true, false, P2, FMX>), dim3((unsigned)grid), dim3(FMX / 16),
                               0, s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, o, ostride, col0, ctr);
        else
            hipLaunchKernelGGL((k_fstage_direct_p<M, T, false, false, P2, FMX>), dim3((unsigned)grid), dim3(FMX / 16),
                               0, s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, o, ostride, col0, ctr);
    } else if (pm && !mult && !sc)
        hipLaunchKernelGGL((k_fstage_direct_p<M, T, true, true, false, FMX>), dim3((unsigned)grid), dim3(FMX / 16), 0, s,
                           tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, o, ostride, col0, ctr);
    else if (pm)
        hipLaunchKernelGGL((k_fstage_direct_p<M, T, true, false, false, FMX>), dim3((unsigned)grid), dim3(FMX / 16), 0, s,
                           tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, o, ostride, col0, ctr);
    else
        hipLaunchKernelGGL((k_fstage_direct_p<M, T, false, false, false, FMX>), dim3((unsigned)grid), dim3(FMX / 16), 0, s,
                           tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, o, ostride, col0, ctr);
    return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP;
}

template <int M>
int launch_fstage_direct(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp,
                         long t_first, long nt, const fx_phase_model* pm, const float2* mult, const float* sc, float su,
                         const float* si, int out_type, void* out, long ostride, long col0, long ostride_i,
                         unsigned* ctr, int nwg, bool pair)
{
    switch (out_type) {
    case FX_PREC_F32:
        if (ctr && nwg >= 8 && ninp <= FX_PDIR_PM)
            return launch_fstage_direct_p<M, float2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si,
                                                     out, ostride, col0, ctr, nwg, pair);
        return launch_fstage_direct_t<M, float2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out,
                                                 ostride, col0, ostride_i);
    case FX_PREC_F16:
        if (ctr && nwg >= 8 && ninp <= FX_PDIR_PM)
            return launch_fstage_direct_p<M, __half2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si,
                                                      out, ostride, col0, ctr, nwg, pair);
        return launch_fstage_direct_t<M, __half2>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out,
                                                  ostride, col0, ostride_i);
    case FX_PREC_FP8:
        return launch_fstage_direct_t<M, uint16_t>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si,
                                                   out, ostride, col0, ostride_i);
    default:
        return FX_ERR_ARG;
    }
}

} // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int fxk_fft_c2c(hipStream_t s, const float2* tw, const float2* in, float2* out, int m, long batch, int inverse,
                float scale)
{
    switch (m) {
    case 2: return launch_c2c<2>(s, tw, in, out, batch, inverse, scale);
    case 4: return launch_c2c<4>(s, tw, in, out, batch, inverse, scale);
    case 8: return launch_c2c<8>(s, tw, in, out, batch, inverse, scale);
    case 16: return launch_c2c<16>(s, tw, in, out, batch, inverse, scale);
    case 32: return launch_c2c<32>(s, tw, in, out, batch, inverse, scale);
    case 64: return launch_c2c<64>(s, tw, in, out, batch, inverse, scale);
    case 128: return launch_c2c<128>(s, tw, in, out, batch, inverse, scale);
    case 256: return launch_c2c<256>(s, tw, in, out, batch, inverse, scale);
    case 512: return launch_c2c<512>(s, tw, in, out, batch, inverse, scale);
    case 1024: return launch_c2c<1024>(s, tw, in, out, batch, inverse, scale);
    case 2048: return launch_c2c<2048>(s, tw, in, out, batch, inverse, scale);
    case 4096: return launch_c2c<4096>(s, tw, in, out, batch, inverse, scale);
    case 8192: return launch_c2c<8192>(s, tw, in, out, batch, inverse, scale);
    case 16384: return launch_c2c<16384>(s, tw, in, out, batch, inverse, scale);
    default: return FX_ERR_ARG;
    }
}

int fxk_fft_r2c(hipStream_t s, const float2* tw, const float* in, float2* out, int n, long batch,
                int keep)
{
    switch (n) {
    case 4: return launch_r2c<2>(s, tw, in, out, batch, keep);
    case 8: return launch_r2c<4>(s, tw, in, out, batch, keep);
    case 16: return launch_r2c<8>(s, tw, in, out, batch, keep);
    case 32: return launch_r2c<16>(s, tw, in, out, batch, keep);
    case 64: return launch_r2c<32>(s, tw, in, out, batch, keep);
    case 128: return launch_r2c<64>(s, tw, in, out, batch, keep);
    case 256: return launch_r2c<128>(s, tw, in, out, batch, keep);
    case 512: return launch_r2c<256>(s, tw, in, out, batch, keep);
    case 1024: return launch_r2c<512>(s, tw, in, out, batch, keep);
    case 2048: return launch_r2c<1024>(s, tw, in, out, batch, keep);
    case 4096: return launch_r2c<2048>(s, tw, in, out, batch, keep);
    case 8192: return launch_r2c<4096>(s, tw, in, out, batch, keep);
    case 16384: return launch_r2c<8192>(s, tw, in, out, batch, keep);
    case 32768: return launch_r2c<16384>(s, tw, in, out, batch, keep);
    default: return FX_ERR_ARG;
    }
}

int fxk_spectrometer(hipStream_t s, const float2* tw, const float* win, const uint32_t* words,
                     long hop, long spec_stride, int nchan, int n, int integrate_ffts, int nspec, float* partial,
                     int groups, int nyq, float* out)
{
    switch (n) {
    case 256: return launch_spec<128>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 512: return launch_spec<256>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 1024: return launch_spec<512>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 2048: return launch_spec<1024>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 4096: return launch_spec<2048>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 8192: return launch_spec<4096>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 16384: return launch_spec<8192>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    case 32768: return launch_spec<16384>(s, tw, win, words, hop, spec_stride, nchan, integrate_ffts, nspec, groups, nyq, partial, out);
    default: return FX_ERR_ARG;
    }
}

int fxk_fstage(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp,
               int n, long t_first, long nt, const fx_phase_model* pm, const float2* mult,
               const float* sc, float su, const float* si, int out_type, void* nat, void* out)
{
    switch (n) {
    case 256: return launch_fstage<128>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 512: return launch_fstage<256>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 1024: return launch_fstage<512>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 2048: return launch_fstage<1024>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 4096: return launch_fstage<2048>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 8192: return launch_fstage<4096>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    case 16384: return launch_fstage<8192>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, nat, out);
    default: return FX_ERR_ARG;
    }
}

int fxk_cornerturn(hipStream_t s, int out_type, const void* nat, void* out, long nitems, int nchan,
                   long ostride, long col0, int ns, long ostride_i)
{
    if (nchan % 64) return FX_ERR_ARG;
    switch (out_type) {
    case FX_PREC_F32:
        return launch_cornerturn_t<float2>(s, nat, out, nitems, nchan, ostride, col0, 0, ns, ostride_i);
    case FX_PREC_F16:
        return launch_cornerturn_t<uint32_t>(s, nat, out, nitems, nchan, ostride, col0, 0, ns, ostride_i);
    case FX_PREC_FP8:
        return launch_cornerturn_t<uint16_t>(s, nat, out, nitems, nchan, ostride, col0, 1, ns, ostride_i);
    default: return FX_ERR_ARG;
    }
}

int fxk_fstage_direct(hipStream_t s, const float2* tw, const uint32_t* words, size_t stride, int ninp, int n,
                      long t_first, long nt, const fx_phase_model* pm, const float2* mult, const float* sc, float su,
                      const float* si, int out_type, void* out, long ostride, long col0, long ostride_i,
                      unsigned* ctr, int nwg, bool pair)
{
    switch (n) {
    case 4096: return launch_fstage_direct<2048>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, out, ostride, col0, ostride_i, ctr, nwg, pair);
    case 8192: return launch_fstage_direct<4096>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, out, ostride, col0, ostride_i, ctr, nwg, pair);
    case 16384: return launch_fstage_direct<8192>(s, tw, words, stride, ninp, t_first, nt, pm, mult, sc, su, si, out_type, out, ostride, col0, ostride_i, ctr, nwg, pair);
    default: return FX_ERR_ARG;
    }
}

bool fxk_fstage_direct_ok(int n, int ninp, long nt, int out_type, long ostride, long col0, long ostride_i)
{
    if (n != 4096 && n != 8192 && n != 16384) return false;
    // columns per workgroup: (points per round / M) x rounds, M = n/2
    const long F = out_type == FX_PREC_F32   ? 2L * direct_fm<float2>() / n * direct_rounds<float2>()
                 : out_type == FX_PREC_F16 ? 2L * direct_fm<__half2>() / n * direct_rounds<__half2>()
                                           : 2L * direct_fm<uint16_t>() / n * direct_rounds<uint16_t>();
    if (out_type == FX_PREC_FP8) return nt % F == 0 && ostride % F == 0 && col0 % F == 0 && ostride_i % F == 0;
    if (out_type != FX_PREC_F32 && out_type != FX_PREC_F16) return false;
    return ninp % F == 0 && ostride % F == 0 && col0 % F == 0;
}

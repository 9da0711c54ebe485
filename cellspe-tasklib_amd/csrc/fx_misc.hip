// fx_misc.hip -- unpack, synthetic generators, phase-cal and elementwise DSP
// kernels of the task library.
//
//   k_unpack2b         2-bit -> float through {-H,-1,+1,+H}
//                      (more_raws_to_complex, minicorrelator_spu.c:662-692;
//                      TwoBitUnpacker::extract_samples, DataUnpackers.cpp:174-227;
//                      VDIF convention format_vdif.c:66-96)
//   k_pcal_fold        exact FP64 fold of 2-bit samples into the phase-cal
//                      period (PCal.cpp:265-288, 424-479, 610-644)
//   k_pcal_final       tone extraction DFT (PCal.cpp:298-314, 489-515, 654-678)
//   k_cmac / k_cmul    genericAddProduct_32fc[_conj] / genericMul_32fc
//                      (architecture.h:270-277, 405-434), do_complex_mac
//                      (cellspe-tasklib-dsp.cpp:454-514)
//   k_sincos*          exec_sincos / exec_sincos_generator
//                      (cellspe-tasklib-dsp.cpp:704-804)
#include "fx_internal.h"

namespace {

// ---------------------------------------------------------------------------
// unpack: thread per packed word; out[c*ntime + t]
// ---------------------------------------------------------------------------
__global__ void k_unpack2b(const uint32_t* __restrict__ words, size_t ntime, int nchan,
                           float* __restrict__ out)
{
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int spw = 16 / nchan;
    const size_t t0 = w * (size_t)spw;
    if (t0 >= ntime) return;
    const uint32_t v = words[w];
    if (nchan == 1) {
        // a wave's 64 words are 1024 samples: store q puts lane l's float4
        // at samples 4 (64 q + l) of that span, so every store is a contiguous
        // 1 KB; the byte it needs is byte l % 4 of word 16 q + l / 4, taken
        // from that lane with a cross-lane read
        const int lane = threadIdx.x & 63;
        const size_t w0 = w - (size_t)lane;
        if ((w0 + 64) * 16 <= ntime) {
            float4* o = reinterpret_cast<float4*>(out + w0 * 16);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t src = (uint32_t)__shfl((int)v, 16 * q + (lane >> 2), 64);
                const uint32_t b = src >> (8 * (lane & 3));
                o[64 * q + lane] = make_float4(lut2(b & 3u), lut2((b >> 2) & 3u), lut2((b >> 4) & 3u),
                                               lut2((b >> 6) & 3u));
            }
            return;
        }
        if (t0 + 16 <= ntime) {
            float4* o = reinterpret_cast<float4*>(out + t0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int b = 8 * q;
                o[q] = make_float4(lut2((v >> b) & 3u), lut2((v >> (b + 2)) & 3u),
                                   lut2((v >> (b + 4)) & 3u), lut2((v >> (b + 6)) & 3u));
            }
            return;
        }
    }
    for (int s = 0; s < spw; s++) {
        size_t t = t0 + (size_t)s;
        if (t >= ntime) break;
        for (int c = 0; c < nchan; c++)
            out[(size_t)c * ntime + t] = lut2((v >> (2 * (s * nchan + c))) & 3u);
    }
}

// ---------------------------------------------------------------------------
// synthetic generators: the oracle's counter-based model evaluated per word
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t key, uint64_t ctr)
{
    uint64_t r = mix64(key ^ mix64(ctr));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double gauss(uint64_t key, uint64_t ctr)
{
    double u1 = uniform01(key, 2 * ctr);
    double u2 = uniform01(key, 2 * ctr + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
__device__ __forceinline__ uint32_t quant2(double x, double t)
{
    return x < -t ? 0u : (x < 0.0 ? 1u : (x < t ? 2u : 3u));
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    auto m = [](uint64_t x) {
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    };
    return m(seed ^ m(stream + 0x51ED270B27A3D1C5ull));
}

__global__ void k_synth_station(uint64_t kc, uint64_t ko, double rho, long delay, double tone_amp,
                                double tone_freq, double thresh, size_t nsamp,
                                uint32_t* __restrict__ words)
{
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 16 >= nsamp) return;
    const double a = sqrt(1.0 - rho * rho);
    uint32_t v = 0;
    for (int j = 0; j < 16; j++) {
        size_t n = w * 16 + (size_t)j;
        if (n >= nsamp) break;
        double x = a * gauss(ko, n);
        if (rho != 0.0) x += rho * gauss(kc, (uint64_t)((long)n - delay + 64));
        if (tone_amp != 0.0) {
            double ph = tone_freq * (double)n;
            ph -= floor(ph);
            x += tone_amp * cos(6.283185307179586 * ph);
        }
        v |= quant2(x, thresh) << (2 * j);
    }
    words[w] = v;
}

struct ToneSet {
    double amp[16];
    double freq[16];
};

__global__ void k_synth_multichannel(uint64_t seed, int nchan, ToneSet tones, double thresh,
                                     size_t ntime, uint32_t* __restrict__ words)
{
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int spw = 16 / nchan;
    if (w * (size_t)spw >= ntime) return;
    uint32_t v = 0;
    for (int s = 0; s < spw; s++) {
        size_t t = w * (size_t)spw + (size_t)s;
        if (t >= ntime) break;
        for (int c = 0; c < nchan; c++) {
            double x = gauss(stream_key(seed, 0x200 + (uint64_t)c), t);
            if (tones.amp[c] != 0.0) {
                double ph = tones.freq[c] * (double)t;
                ph -= floor(ph);
                x += tones.amp[c] * cos(6.283185307179586 * ph);
            }
            v |= quant2(x, thresh) << (2 * (s * nchan + c));
        }
    }
    words[w] = v;
}

__global__ void k_synth_channelized(uint64_t seed, int nchan, long ntime, int ninp, double rho,
                                    double thresh, int imajor, uint8_t* __restrict__ codes)
{
    const size_t total = (size_t)nchan * (size_t)ntime * (size_t)ninp;
    // grid-stride: totals beyond 2^32 work items exceed one launch's grid
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i, c;
        long t;
        if (imajor) {   // codes[c][i][t]
            t = (long)(idx % (size_t)ntime);
            const size_t ci = idx / (size_t)ntime;
            i = (int)(ci % (size_t)ninp);
            c = (int)(ci / (size_t)ninp);
        } else {        // codes[c][t][i]
            i = (int)(idx % (size_t)ninp);
            const size_t ct = idx / (size_t)ninp;
            t = (long)(ct % (size_t)ntime);
            c = (int)(ct / (size_t)ntime);
        }
        const double a = sqrt(1.0 - rho * rho);
        const uint64_t kc = stream_key(seed, 0x10000 + (uint64_t)c);
        const uint64_t ko = stream_key(seed, 0x100000000ull + (uint64_t)c * 65536u + (uint64_t)i);
        double xr = rho * gauss(kc, (uint64_t)(2 * t)) + a * gauss(ko, (uint64_t)(2 * t));
        double xi = rho * gauss(kc, (uint64_t)(2 * t + 1)) + a * gauss(ko, (uint64_t)(2 * t + 1));
        codes[idx] = (uint8_t)(quant2(xr, thresh) | (quant2(xi, thresh) << 2));
    }
}

// ---------------------------------------------------------------------------
// phase-cal fold.  acc[(position + n) mod nfold] += LUT(code_n) in FP64.  The
// summands are FP32 level values (multiples of 2^-22, |x| < 4), so every FP64
// partial sum below 2^31 is exact: the result does not depend on the order of
// the atomic additions and is bit-identical run to run.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pcal_fold(const uint32_t* __restrict__ words, size_t nsamp,
                                                   size_t position, int nfold, int use_lds,
                                                   double* __restrict__ acc, size_t in_stride,
                                                   const uint8_t* __restrict__ wvalid)
{
    extern __shared__ double lacc[];
    // grid y = input: words of input y at words + y*in_stride, its fold at acc + y*nfold
    words += (size_t)blockIdx.y * in_stride;
    acc += (size_t)blockIdx.y * (size_t)nfold;
    if (use_lds) {
        for (int b = threadIdx.x; b < nfold; b += blockDim.x) lacc[b] = 0.0;
        __syncthreads();
    }
    // grid-stride over the packed words: a workgroup folds many words into its
    // LDS copy before the nfold global additions of its partial fold
    const size_t nw = (nsamp + 15) / 16;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (size_t)gridDim.x * blockDim.x) {
        if (wvalid && !wvalid[w]) continue;         // a blanked word decodes as zeros
        const uint32_t v = words[w];
        size_t n0 = w * 16;
        int bin = (int)((position + n0) % (size_t)nfold);
        for (int j = 0; j < 16; j++) {
            if (n0 + (size_t)j >= nsamp) break;
            double x = (double)lut2((v >> (2 * j)) & 3u);
            if (use_lds)
                atomicAdd(&lacc[bin], x);
            else
                atomicAdd(&acc[bin], x);
            if (++bin == nfold) bin = 0;
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int b = threadIdx.x; b < nfold; b += blockDim.x)
            if (lacc[b] != 0.0) atomicAdd(&acc[b], lacc[b]);
    }
}

// nfold % 16 == 0: a word's 16 samples land in 16 consecutive bins, and with a
// word stride T that is a multiple of nfold/16 every thread keeps the same 16
// bins for all its words.  Each thread counts codes per bin in registers --
// A = n(+H) - n(-H), B = n(+1) - n(-1), so the bin's sum is H A + B, four
// bins per 32-bit add in byte lanes (Pcal16) -- and adds its 16 sums once.  The FP64 sums are the same exact values as the
// sample-by-sample fold (multiples of 2^-22 below 2^31).
__global__ __launch_bounds__(256) void k_pcal_fold16(const uint32_t* __restrict__ words, size_t nsamp,
                                                     size_t position, int nfold, long T,
                                                     double* __restrict__ acc, size_t in_stride,
                                                     const uint8_t* __restrict__ wvalid)
{
    extern __shared__ double lacc[];
    words += (size_t)blockIdx.y * in_stride;
    acc += (size_t)blockIdx.y * (size_t)nfold;
    for (int b = threadIdx.x; b < nfold; b += blockDim.x) lacc[b] = 0.0;
    __syncthreads();
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nw = (nsamp + 15) / 16;
    const size_t nfull = nsamp / 16;   // words whose 16 samples all count
    Pcal16 cnt;
    cnt.init();
    if (tid < T) {
        for (size_t w = (size_t)tid; w < nw; w += (size_t)T) {
            if (wvalid && !wvalid[w]) continue;     // a blanked word decodes as zeros
            const uint32_t v = words[w];
            if (w < nfull)
                cnt.add(v);
            else
                cnt.add_partial(v, (int)(nsamp - w * 16));
        }
        cnt.flush();
        int bin = (int)((position + 16 * (size_t)tid) % (size_t)nfold);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (cnt.A[j] | cnt.B[j])
                atomicAdd(&lacc[bin], (double)FX_LUT_HI * (double)cnt.A[j] + (double)cnt.B[j]);
            if (++bin == nfold) bin = 0;
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nfold; b += blockDim.x)
        if (lacc[b] != 0.0) atomicAdd(&acc[b], lacc[b]);
}

// tone t at DFT bin k = offset_bin + t*step of the nbins-point fold:
//   A[b] = sum_{r = b mod nbins} R[r] e^{i dphi (r mod rotatorlen)}
//          (Shifting extractor, nfold = lcm(rotatorlen, nbins))
//   A[b] = R[b]  (rotatorlen == 0: Trivial / ImplicitShift, nfold = nbins)
//   tone = (1/nbins) sum_b A[b] e^{-2 pi i b k/nbins}
// One workgroup per tone: the threads share the fold bins and their partial
// sums are reduced through LDS in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_pcal_final(const double* __restrict__ acc, int nfold, int nbins,
                                                    int ntones, int offset_bin, int step, int rotatorlen,
                                                    double dphi, double* __restrict__ tones)
{
    __shared__ double red[2][256];
    const int t = blockIdx.x;
    if (t >= ntones) return;
    // grid y = input
    acc += (size_t)blockIdx.y * (size_t)nfold;
    tones += (size_t)blockIdx.y * 2 * (size_t)ntones;
    const long k = (long)offset_bin + (long)t * step;
    double sr = 0.0, si = 0.0;
    if (k < nbins) {
        for (int r = threadIdx.x; r < nfold; r += blockDim.x) {
            double v = acc[r];
            if (v == 0.0) continue;
            long e = ((long)(r % nbins) * k) % nbins;
            double ang = -6.283185307179586 * (double)e / (double)nbins;
            if (rotatorlen) ang += dphi * (double)(r % rotatorlen);
            double sn, cs;
            sincos(ang, &sn, &cs);
            sr += v * cs;
            si += v * sn;
        }
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tones[2 * t] = k < nbins ? red[0][0] / (double)nbins : 0.0;
        tones[2 * t + 1] = k < nbins ? red[1][0] / (double)nbins : 0.0;
    }
}

// ---------------------------------------------------------------------------
// elementwise DSP
// ---------------------------------------------------------------------------
__global__ void k_cmac(const float2* __restrict__ a, const float2* __restrict__ b,
                       float2* __restrict__ acc, long n, int conj)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 x = a[i], y = b[i], s = acc[i];
    if (conj) {
        s.x += (x.x * y.x) + (x.y * y.y);
        s.y += (x.y * y.x) - (x.x * y.y);
    } else {
        s.x += (x.x * y.x) - (x.y * y.y);
        s.y += (x.y * y.x) + (x.x * y.y);
    }
    acc[i] = s;
}

__global__ void k_cmul(const float2* __restrict__ a, const float2* __restrict__ b,
                       float2* __restrict__ out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 x = a[i], y = b[i];
    out[i] = make_float2((x.x * y.x) - (x.y * y.y), (x.y * y.x) + (x.x * y.y));
}

__global__ void k_sincos(const float* __restrict__ arg, float2* __restrict__ out, long n, int norm)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s, c;
    if (norm) {
        // turns: reduce in FP64 so large arguments keep their fraction
        double cyc = (double)arg[i];
        cyc -= rint(cyc);
        sincosf((float)(6.283185307179586 * cyc), &s, &c);
    } else {
        sincosf(arg[i], &s, &c);
    }
    out[i] = make_float2(c, s);
}

__global__ void k_sincos_gen(double start, double inc, float2* __restrict__ out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ph = start + inc * (double)i;
    double cyc = ph * 0.15915494309189535;
    cyc -= rint(cyc);
    float s, c;
    sincosf((float)(6.283185307179586 * cyc), &s, &c);
    out[i] = make_float2(c, s);
}

// plain 16-byte-per-lane copy (traffic calibration and staging)
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_cadd(const float2* __restrict__ a, const float2* __restrict__ b,
                       float2* __restrict__ out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 x = a[i], y = b[i];
    out[i] = make_float2(x.x + y.x, x.y + y.y);
}

// amplitude and phase of scaled visibilities or lags, in FP64 then rounded
// (zerocorr.c:628-630,639-643: x, y scaled by 1/n_ffts, sqrt(x^2+y^2),
// atan2(y, x)); grid-stride
__global__ void k_amp_phase(const float2* __restrict__ z, long n, float scale, float* __restrict__ amp,
                            float* __restrict__ phase)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float2 v = z[i];
        const double x = (double)v.x * (double)scale, y = (double)v.y * (double)scale;
        if (amp) amp[i] = (float)sqrt(x * x + y * y);
        if (phase) phase[i] = (float)atan2(y, x);
    }
}

// Phase-cal delay of each input (m5pcal.c:94-150, calcDelay): tone phases
// phi_t = atan2(im, re) of tones at f_t = f0 + t df; reference slope from the
// middle pair (ntones/2, ntones/2 + 1), or (0, 1) for two tones (the
// reference reads past the end there); every adjacent pair with both tones
// within band_valid of band_center contributes (phi_t - phi_{t-1}), folded
// once by 2 pi toward the reference slope, over (f_t - f_{t-1}) in MHz; the
// mean slope is turned into ns (1000 / 2 pi) and wrapped into
// +-500 / (tone spacing in MHz).  One thread per input.
__global__ void k_pcal_delay(const double* __restrict__ tones, int ninp, int ntones, double f0_mhz,
                             double df_mhz, double band_center_mhz, double band_valid_mhz,
                             double* __restrict__ delay_ns)
{
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ninp) return;
    const double* z = tones + (size_t)y * 2 * (size_t)ntones;
    if (ntones < 2) {
        delay_ns[y] = 0.0;
        return;
    }
    const double pi = 3.141592653589793;
    const int a = ntones >= 3 ? ntones / 2 : 0;
    const double ref = atan2(z[2 * a + 3], z[2 * a + 2]) - atan2(z[2 * a + 1], z[2 * a]);
    double sum = 0.0;
    int n = 0;
    double prev = atan2(z[1], z[0]);
    for (int t = 1; t < ntones; t++) {
        const double ph = atan2(z[2 * t + 1], z[2 * t]);
        const double fa = f0_mhz + (t - 1) * df_mhz, fb = f0_mhz + t * df_mhz;
        if (fabs(fa - band_center_mhz) <= band_valid_mhz && fabs(fb - band_center_mhz) <= band_valid_mhz) {
            double d = ph - prev;
            if (d - ref > pi) d -= 2.0 * pi;
            if (ref - d > pi) d += 2.0 * pi;
            sum += d / (fb - fa);
            n++;
        }
        prev = ph;
    }
    if (n < 1) {
        delay_ns[y] = 0.0;
        return;
    }
    const double window = fabs(500.0 / df_mhz);
    double delay = 1000.0 * (sum / n) / (2.0 * pi);
    while (delay > window) delay -= 2.0 * window;
    while (delay < -window) delay += 2.0 * window;
    delay_ns[y] = delay;
}

// DiFX-process epilogue: y = band * mult; out = y; conj_out = conj(y);
// acc.re += |y|^2 (acc.im untouched, autocorrelations are real)
__global__ void k_difx_post(const float2* __restrict__ fft, int nfft, int band0, int n,
                            const float2* __restrict__ mult, float2* __restrict__ out,
                            float2* __restrict__ conj_out, float2* __restrict__ acc)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float2 x = fft[(band0 + k) & (nfft - 1)], m = mult[k];
    float2 y = make_float2((x.x * m.x) - (x.y * m.y), (x.y * m.x) + (x.x * m.y));
    out[k] = y;
    conj_out[k] = make_float2(y.x, -y.y);
    acc[k].x += y.x * y.x + y.y * y.y;
}

// grid-stride 32-bit fill (clears that may be captured in a graph)
__global__ void k_fill32(uint32_t* __restrict__ p, uint32_t value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Spectrometer window table w[i], i < n, evaluated in FP64 and rounded once:
// SIN2 sin^2(pi i/(n-1)) (spectrometer_spu.c:297-315); COSINE and COSINE2
// over n points centred between samples, and the generalized cosine
// windows Hamming, Blackman (alpha 0.5) and Hann over n-1 intervals
// (TaskCoreIPP.cpp:569-605); NONE all ones.
__global__ void k_window(int kind, int n, float* __restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pi = 3.14159265358979323846;
    double v = 1.0;
    if (kind == FX_WIN_SIN2) {
        const double s = sin(pi * (double)i / (double)(n - 1));
        v = s * s;
    } else if (kind == FX_WIN_COSINE) {
        v = cos(pi * ((double)i - 0.5 * (double)n + 0.5) / (double)n);
    } else if (kind == FX_WIN_COSINE2) {
        const double c = cos(pi * ((double)i - 0.5 * (double)n + 0.5) / (double)n);
        v = c * c;
    } else if (kind >= FX_WIN_HAMMING) {
        const double x = 2.0 * pi * (double)i / (double)(n - 1);
        if (kind == FX_WIN_HAMMING)
            v = 0.54 - 0.46 * cos(x);
        else if (kind == FX_WIN_BLACKMAN)
            v = 0.75 - 0.5 * cos(x) - 0.25 * cos(2.0 * x);
        else
            v = 0.5 - 0.5 * cos(x);
    }
    w[i] = (float)v;
}

__global__ void k_iotest(float* out)
{
    if (threadIdx.x == 0) {
        out[0] = 0.1f;
        out[1] = 2.3f;
        out[2] = 4.5f;
        out[3] = 6.7f;
    }
}

inline unsigned nblocks(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }
// grid for the grid-stride kernels: enough workgroups to fill the chip
inline unsigned nblocks_gs(size_t n, unsigned bs) { return (unsigned)std::min<size_t>((n + bs - 1) / bs, 1u << 18); }
inline int last_status() { return hipGetLastError() == hipSuccess ? FX_OK : FX_ERR_HIP; }

} // namespace

int fxk_unpack2b(hipStream_t s, const uint32_t* words, size_t ntime, int nchan, float* out)
{
    size_t nw = (ntime * (size_t)nchan + 15) / 16;
    if (nw == 0) return FX_OK;
    hipLaunchKernelGGL(k_unpack2b, dim3(nblocks(nw, 256)), dim3(256), 0, s, words, ntime, nchan, out);
    return last_status();
}

int fxk_synth_station(hipStream_t s, uint64_t seed, int station, double rho, long delay,
                      double tone_amp, double tone_freq, double thresh, size_t nsamp,
                      uint32_t* words)
{
    size_t nw = (nsamp + 15) / 16;
    if (nw == 0) return FX_OK;
    uint64_t kc = stream_key(seed, 0xC0);
    uint64_t ko = stream_key(seed, 0x100 + (uint64_t)station);
    hipLaunchKernelGGL(k_synth_station, dim3(nblocks(nw, 256)), dim3(256), 0, s, kc, ko, rho, delay,
                       tone_amp, tone_freq, thresh, nsamp, words);
    return last_status();
}

int fxk_synth_multichannel(hipStream_t s, uint64_t seed, int nchan, const double* tone_amp,
                           const double* tone_freq, double thresh, size_t ntime, uint32_t* words)
{
    ToneSet ts;
    for (int c = 0; c < 16; c++) {
        ts.amp[c] = (tone_amp && c < nchan) ? tone_amp[c] : 0.0;
        ts.freq[c] = (tone_freq && c < nchan) ? tone_freq[c] : 0.0;
    }
    int spw = 16 / nchan;
    size_t nw = (ntime + (size_t)spw - 1) / (size_t)spw;
    if (nw == 0) return FX_OK;
    hipLaunchKernelGGL(k_synth_multichannel, dim3(nblocks(nw, 256)), dim3(256), 0, s, seed, nchan,
                       ts, thresh, ntime, words);
    return last_status();
}

int fxk_synth_channelized(hipStream_t s, uint64_t seed, int nchan, long ntime, int ninp,
                          double rho, double thresh, int imajor, uint8_t* codes)
{
    size_t total = (size_t)nchan * (size_t)ntime * (size_t)ninp;
    if (total == 0) return FX_OK;
    hipLaunchKernelGGL(k_synth_channelized, dim3(nblocks_gs(total, 256)), dim3(256), 0, s, seed, nchan,
                       ntime, ninp, rho, thresh, imajor, codes);
    return last_status();
}

// 2-bit channelized codes -> [c][t][i] FP32 / FP16 samples through the level
// table (the FP32 and FP16 engines' input for channelized data)
template <typename V>
__global__ void k_codes_to_chan(const uint8_t* __restrict__ codes, int in_imajor, int nchan, long nt,
                                int ninp, V* __restrict__ out)
{
    const size_t total = (size_t)nchan * (size_t)nt * (size_t)ninp;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % (size_t)ninp);
        const size_t ct = idx / (size_t)ninp;
        const long t = (long)(ct % (size_t)nt);
        const int c = (int)(ct / (size_t)nt);
        const uint32_t code = in_imajor ? codes[((size_t)c * ninp + i) * (size_t)nt + t] : codes[idx];
        const float re = lut2(code & 3u), im = lut2((code >> 2) & 3u);
        if constexpr (sizeof(V) == 8) {
            out[idx] = make_float2(re, im);
        } else {
            __half2 h = __floats2half2_rn(re, im);
            out[idx] = *reinterpret_cast<uint32_t*>(&h);
        }
    }
}

int fxk_codes_to_chan(hipStream_t s, const uint8_t* codes, int in_imajor, int nchan, long nt, int ninp,
                      int out_prec, void* out)
{
    size_t total = (size_t)nchan * (size_t)nt * (size_t)ninp;
    if (total == 0) return FX_OK;
    if (out_prec == FX_PREC_F32)
        hipLaunchKernelGGL(k_codes_to_chan<float2>, dim3(nblocks_gs(total, 256)), dim3(256), 0, s, codes,
                           in_imajor, nchan, nt, ninp, reinterpret_cast<float2*>(out));
    else if (out_prec == FX_PREC_F16)
        hipLaunchKernelGGL(k_codes_to_chan<uint32_t>, dim3(nblocks_gs(total, 256)), dim3(256), 0, s, codes,
                           in_imajor, nchan, nt, ninp, reinterpret_cast<uint32_t*>(out));
    else
        return FX_ERR_ARG;
    return last_status();
}

int fxk_pcal_fold(hipStream_t s, const uint32_t* words, size_t nsamp, size_t position, int nfold,
                  double* acc, int ninp, size_t in_stride_words, const uint8_t* wvalid)
{
    size_t nw = (nsamp + 15) / 16;
    if (nw == 0 || ninp <= 0) return FX_OK;
    if (nfold % 16 == 0 && nfold <= 8192) {
        // word stride: a multiple of nfold/16, about 8+ words per thread
        const long nb = nfold / 16;
        const long want = std::max<long>(256, std::min<long>(65536, (long)(nw / 8)));
        const long T = std::max<long>(1, want / nb) * nb;
        const unsigned blocks = (unsigned)((T + 255) / 256);
        hipLaunchKernelGGL(k_pcal_fold16, dim3(blocks, (unsigned)ninp), dim3(256), (size_t)nfold * sizeof(double), s,
                           words, nsamp, position, nfold, T, acc, in_stride_words, wvalid);
        return last_status();
    }
    int use_lds = nfold <= 8192;
    size_t lds = use_lds ? (size_t)nfold * sizeof(double) : 0;
    // at least 16 words per thread: each workgroup's nfold global additions
    // are amortized over >= 65536 samples
    const size_t blocks = std::min<size_t>(nblocks(nw, 256), std::max<size_t>(64, nw / 4096));
    hipLaunchKernelGGL(k_pcal_fold, dim3((unsigned)blocks, (unsigned)ninp), dim3(256), lds, s, words, nsamp,
                       position, nfold, use_lds, acc, in_stride_words, wvalid);
    return last_status();
}

int fxk_pcal_final(hipStream_t s, const double* acc, int nfold, int nbins, int ntones,
                   int offset_bin, int step, int rotatorlen, double dphi, double* tones, int ninp)
{
    if (ntones <= 0 || ninp <= 0) return FX_OK;
    hipLaunchKernelGGL(k_pcal_final, dim3((unsigned)ntones, (unsigned)ninp), dim3(256), 0, s, acc, nfold, nbins,
                       ntones, offset_bin, step, rotatorlen, dphi, tones);
    return last_status();
}

int fxk_complex_mac(hipStream_t s, const float2* a, const float2* b, float2* acc, long n, int conj)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_cmac, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, a, b, acc, n, conj);
    return last_status();
}

int fxk_complex_mul(hipStream_t s, const float2* a, const float2* b, float2* out, long n)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_cmul, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, a, b, out, n);
    return last_status();
}

int fxk_sincos(hipStream_t s, const float* arg, float2* out, long n, int normalized)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_sincos, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, arg, out, n,
                       normalized);
    return last_status();
}

int fxk_sincos_gen(hipStream_t s, double start, double inc, float2* out, long n)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_sincos_gen, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, start, inc,
                       out, n);
    return last_status();
}

int fxk_copy(hipStream_t s, const void* src, void* dst, size_t bytes)
{
    if (bytes % 16 || ((uintptr_t)src % 16) || ((uintptr_t)dst % 16)) return FX_ERR_ARG;
    size_t n16 = bytes / 16;
    if (n16 == 0) return FX_OK;
    unsigned blocks = (unsigned)std::min<size_t>((n16 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
    return last_status();
}

int fxk_complex_add(hipStream_t s, const float2* a, const float2* b, float2* out, long n)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_cadd, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, a, b, out, n);
    return last_status();
}

int fxk_amp_phase(hipStream_t s, const float2* z, long n, float scale, float* amp, float* phase)
{
    if (n <= 0) return FX_OK;
    const size_t blocks = std::min<size_t>(nblocks((size_t)n, 256), (size_t)1 << 18);
    hipLaunchKernelGGL(k_amp_phase, dim3((unsigned)blocks), dim3(256), 0, s, z, n, scale, amp, phase);
    return last_status();
}

int fxk_pcal_delay(hipStream_t s, const double* tones, int ninp, int ntones, double f0_hz, double df_hz,
                   double band_center_hz, double band_valid_hz, double* delay_ns)
{
    if (ninp <= 0) return FX_OK;
    hipLaunchKernelGGL(k_pcal_delay, dim3((unsigned)nblocks((size_t)ninp, 64)), dim3(64), 0, s, tones, ninp,
                       ntones, f0_hz * 1e-6, df_hz * 1e-6, band_center_hz * 1e-6, band_valid_hz * 1e-6,
                       delay_ns);
    return last_status();
}

int fxk_difx_post(hipStream_t s, const float2* fft, int nfft, int band0, int n, const float2* mult,
                  float2* out, float2* conj_out, float2* acc)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_difx_post, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, fft, nfft, band0, n,
                       mult, out, conj_out, acc);
    return last_status();
}

int fxk_fill32(hipStream_t s, void* p, uint32_t value, size_t n)
{
    if (n == 0) return FX_OK;
    hipLaunchKernelGGL(k_fill32, dim3(nblocks_gs(n, 256)), dim3(256), 0, s, reinterpret_cast<uint32_t*>(p), value,
                       n);
    return last_status();
}

int fxk_window(hipStream_t s, int kind, int n, float* w)
{
    if (n <= 0) return FX_OK;
    hipLaunchKernelGGL(k_window, dim3(nblocks((size_t)n, 256)), dim3(256), 0, s, kind, n, w);
    return last_status();
}

int fxk_iotest(hipStream_t s, float* out)
{
    hipLaunchKernelGGL(k_iotest, dim3(1), dim3(64), 0, s, out);
    return last_status();
}

// ---------------------------------------------------------------------------
// spectrometer cross spectra (TaskCoreIPP.cpp:456-487)
// ---------------------------------------------------------------------------
namespace {

// Windowed real FFT inputs of one channel of a packed 2-bit stream: frame
// g0 + blockIdx.y = (spectrum g / K, FFT g % K) starts at sample
// (g / K) * spec_stride + (g % K) * hop.
__global__ __launch_bounds__(256) void k_spec_frames(const uint32_t* __restrict__ words,
                                                     const uint8_t* __restrict__ wvalid, int nchan, int ch,
                                                     long hop, long spec_stride, int K, int n, long g0,
                                                     const float* __restrict__ win, float* __restrict__ out)
{
    const long g = g0 + blockIdx.y;
    const long t0 = (g / K) * spec_stride + (g % K) * hop;
    float* o = out + (size_t)blockIdx.y * (size_t)n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const size_t field = (size_t)(t0 + i) * (size_t)nchan + (size_t)ch;
        const uint32_t code = (words[field >> 4] >> (2u * (uint32_t)(field & 15u))) & 3u;
        // a word the blanker marked invalid decodes as 0 (mark5access zeros)
        o[i] = (wvalid && !wvalid[field >> 4]) ? 0.f : lut2(code) * win[i];
    }
}

// Partial sums of A conj(B) per bin of n/2+1 over the FFTs j0..j1-1 of each
// spectrum, j0 = blockIdx.y * per (the FFTs of a spectrum are split over
// gridDim.y groups so that long integrations still fill the chip); DC and
// Nyquist are real products of the real parts (the IPP core keeps them
// apart).  part[g][sp][k].
__global__ __launch_bounds__(256) void k_xpol_part(const float2* __restrict__ A, const float2* __restrict__ B,
                                                   int c1, int K, int per, long nspec, float2* __restrict__ part)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nspec * (long)c1) return;
    const long sp = idx / c1;
    const int k = (int)(idx % c1);
    const bool edge = k == 0 || k == c1 - 1;
    const int j0 = blockIdx.y * per, j1 = min(K, j0 + per);
    float re = 0.f, im = 0.f;
    for (int j = j0; j < j1; j++) {
        const size_t at = ((size_t)(sp * K + j)) * (size_t)c1 + (size_t)k;
        const float2 a = A[at], b = B[at];
        if (edge) {
            re += a.x * b.x;
        } else {
            re += a.x * b.x + a.y * b.y;
            im += a.y * b.x - a.x * b.y;
        }
    }
    part[(size_t)blockIdx.y * (size_t)(nspec * c1) + (size_t)idx] = make_float2(re, im);
}

// out = scale * sum over the groups, in group order (deterministic)
__global__ __launch_bounds__(256) void k_xpol_sum(const float2* __restrict__ part, int groups, long total,
                                                  float scale, float2* __restrict__ out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    float re = 0.f, im = 0.f;
    for (int g = 0; g < groups; g++) {
        const float2 v = part[(size_t)g * (size_t)total + (size_t)idx];
        re += v.x;
        im += v.y;
    }
    out[idx] = make_float2(re * scale, im * scale);
}

} // namespace

int fxk_spec_frames(hipStream_t s, const uint32_t* words, const uint8_t* wvalid, int nchan, int ch, long hop,
                    long spec_stride, int K, int n, long g0, long nframes, const float* win, float* out)
{
    if (nframes <= 0) return FX_OK;
    const int bx = std::min(8, (n + 255) / 256);
    hipLaunchKernelGGL(k_spec_frames, dim3((unsigned)bx, (unsigned)nframes), dim3(256), 0, s, words, wvalid, nchan, ch,
                       hop, spec_stride, K, n, g0, win, out);
    return last_status();
}

int fxk_xpol_groups(int c1, int K, long nspec)
{
    // about 2^18 threads of partial sums, at least 8 FFTs per group
    const long want = (262144 + nspec * (long)c1 - 1) / (nspec * (long)c1);
    return (int)std::max<long>(1, std::min<long>({want, (long)(K + 7) / 8, 65535}));
}

int fxk_xpol_acc(hipStream_t s, const float2* A, const float2* B, int c1, int K, long nspec, float scale,
                 float2* part, int groups, float2* out)
{
    const long total = nspec * (long)c1;
    if (total <= 0) return FX_OK;
    const int per = (K + groups - 1) / groups;
    hipLaunchKernelGGL(k_xpol_part, dim3((unsigned)((total + 255) / 256), (unsigned)groups), dim3(256), 0, s, A, B,
                       c1, K, per, nspec, part);
    if (hipGetLastError() != hipSuccess) return FX_ERR_HIP;
    hipLaunchKernelGGL(k_xpol_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, part, groups, total, scale,
                       out);
    return last_status();
}

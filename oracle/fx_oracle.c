/*
 * fx_oracle.c -- CPU restatement of the cellspe-tasklib F-X chain.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by the smoke()
 * entry of __graft_entry__.py and by the cpu_baseline leg of bench.py.  The
 * HIP product never links or calls this file.
 *
 * Semantics follow the reference's GENERIC C primitives (architecture.h
 * lines 207-613) composed in the order the minicorrelator and the DiFX-process
 * task use them (minicorrelator/minicorrelator_spu.c:326-481,
 * cellspe-tasklib-dsp.cpp:609-698), with the radix-2 FFT of
 * minicorrelator/fft_1d_r2_modified.h:16-76 in place of FFTW (absent here).
 * The SPU code's defects (its FFT is compiled out, the LUT is a placeholder,
 * the X-engine does not conjugate) are NOT reproduced.
 */
#include "fx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ================================================================== */
/* Synthetic signals                                                  */
/* ================================================================== */

/* splitmix64 finalizer: a bijective 64-bit mixer (Steele et al., 2014). */
uint64_t fxo_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t fxo_stream_key(uint64_t seed, uint64_t stream)
{
    return fxo_mix64(seed ^ fxo_mix64(stream + 0x51ED270B27A3D1C5ull));
}

double fxo_uniform(uint64_t key, uint64_t ctr)
{
    uint64_t r = fxo_mix64(key ^ fxo_mix64(ctr));
    /* 53-bit mantissa, offset by half an ulp so the result is in (0,1) */
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

double fxo_gauss(uint64_t key, uint64_t ctr)
{
    double u1 = fxo_uniform(key, 2 * ctr);
    double u2 = fxo_uniform(key, 2 * ctr + 1);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* 4-level quantizer with thresholds {-t, 0, +t}; offset-binary codes so that
 * code c maps to fxo_lut2bit[c] = {-H, -1, +1, +H} (format_vdif.c:70). */
int fxo_quantize2(double x, double t)
{
    if (x < -t) return 0;
    if (x < 0.0) return 1;
    if (x < t) return 2;
    return 3;
}

/* 16 codes per little-endian 32-bit word, sample j at bits 2j+1:2j
 * (VDIF 1-channel 2-bit layout, format_vdif.c:91-95 read byte-wise). */
void fxo_pack2b(const uint8_t* codes, size_t n, uint32_t* words)
{
    size_t nw = (n + 15) / 16;
    for (size_t w = 0; w < nw; w++) {
        uint32_t v = 0;
        for (int j = 0; j < 16; j++) {
            size_t s = w * 16 + (size_t)j;
            if (s < n) v |= (uint32_t)(codes[s] & 3u) << (2 * j);
        }
        words[w] = v;
    }
}

void fxo_gen_station(uint64_t seed, int station, double rho, long delay,
                     double tone_amp, double tone_freq_norm, double thresh,
                     size_t nsamp, uint32_t* words)
{
    uint64_t kc = fxo_stream_key(seed, 0xC0);
    uint64_t ko = fxo_stream_key(seed, 0x100 + (uint64_t)station);
    double a = sqrt(1.0 - rho * rho);
    size_t nw = (nsamp + 15) / 16;
    for (size_t w = 0; w < nw; w++) {
        uint32_t v = 0;
        for (int j = 0; j < 16; j++) {
            size_t n = w * 16 + (size_t)j;
            if (n >= nsamp) break;
            /* common[] is indexed from -64 so delays up to 64 stay in range */
            double c = fxo_gauss(kc, (uint64_t)((long)n - delay + 64));
            double x = rho * c + a * fxo_gauss(ko, n);
            if (tone_amp != 0.0) {
                double ph = tone_freq_norm * (double)n;
                ph -= floor(ph);
                x += tone_amp * cos(2.0 * M_PI * ph);
            }
            v |= (uint32_t)fxo_quantize2(x, thresh) << (2 * j);
        }
        words[w] = v;
    }
}

void fxo_gen_multichannel(uint64_t seed, int nchan, const double* tone_amp,
                          const double* tone_freq_norm, double thresh,
                          size_t ntime, uint32_t* words)
{
    int spw = 16 / nchan;
    size_t nw = (ntime + (size_t)spw - 1) / (size_t)spw;
    memset(words, 0, nw * sizeof(uint32_t));
    for (int c = 0; c < nchan; c++) {
        uint64_t k = fxo_stream_key(seed, 0x200 + (uint64_t)c);
        for (size_t t = 0; t < ntime; t++) {
            double x = fxo_gauss(k, t);
            if (tone_amp && tone_amp[c] != 0.0) {
                double ph = tone_freq_norm[c] * (double)t;
                ph -= floor(ph);
                x += tone_amp[c] * cos(2.0 * M_PI * ph);
            }
            uint32_t code = (uint32_t)fxo_quantize2(x, thresh);
            size_t w = t / (size_t)spw;
            int pos = (int)(t % (size_t)spw) * nchan + c;
            words[w] |= code << (2 * pos);
        }
    }
}

void fxo_gen_channelized_codes(uint64_t seed, int nchan, long ntime, int ninp,
                               double rho, double thresh, uint8_t* codes)
{
    double a = sqrt(1.0 - rho * rho);
    for (int c = 0; c < nchan; c++) {
        uint64_t kc = fxo_stream_key(seed, 0x10000 + (uint64_t)c);
        for (long t = 0; t < ntime; t++) {
            double cre = fxo_gauss(kc, (uint64_t)(2 * t));
            double cim = fxo_gauss(kc, (uint64_t)(2 * t + 1));
            for (int i = 0; i < ninp; i++) {
                uint64_t ko = fxo_stream_key(seed, 0x100000000ull + (uint64_t)c * 65536u + (uint64_t)i);
                double xr = rho * cre + a * fxo_gauss(ko, (uint64_t)(2 * t));
                double xi = rho * cim + a * fxo_gauss(ko, (uint64_t)(2 * t + 1));
                codes[((size_t)c * (size_t)ntime + (size_t)t) * (size_t)ninp + (size_t)i] =
                    (uint8_t)(fxo_quantize2(xr, thresh) | (fxo_quantize2(xi, thresh) << 2));
            }
        }
    }
}

/* ================================================================== */
/* a1: 2-bit unpack                                                   */
/* ================================================================== */

/* mark5access VDIF 4-level table (format_vdif.c:70) with
 * HiMag = OPTIMAL_2BIT_HIGH = 3.3359 (mark5_stream.h:74). */
const float fxo_lut2bit[4] = { -3.3359f, -1.0f, 1.0f, 3.3359f };

/* 1-channel: sample s at bits 2(s%16)+1:2(s%16) of word s/16
 * (format_vdif.c:558-600 decodes byte b's sample i from bits 2i of the byte,
 * bytes in little-endian word order). */
void fxo_unpack2b_1ch(const uint32_t* words, size_t nsamp, float* out)
{
    for (size_t s = 0; s < nsamp; s++)
        out[s] = fxo_lut2bit[(words[s >> 4] >> (2 * (s & 15))) & 3u];
}

/* nchan-channel interleave, out[c*ntime + t] (format_vdif.c:753-800 for the
 * 16-channel case: byte b carries channels 4b..4b+3). */
void fxo_unpack2b_nch(const uint32_t* words, size_t ntime, int nchan, float* out)
{
    int spw = 16 / nchan;
    for (size_t t = 0; t < ntime; t++) {
        uint32_t w = words[t / (size_t)spw];
        int base = (int)(t % (size_t)spw) * nchan;
        for (int c = 0; c < nchan; c++)
            out[(size_t)c * ntime + t] = fxo_lut2bit[(w >> (2 * (base + c))) & 3u];
    }
}

/* ================================================================== */
/* a2/a3: radix-2 decimation-in-time FFT                               */
/* ================================================================== */

/* Twiddle table of n/4 entries, W[i] = cos(2*pi*i/n) - i*sin(2*pi*i/n),
 * built the way init_FFT does it (cellspe-tasklib-dsp.cpp:531-539): the
 * cosine is evaluated in double and rounded to float, and the imaginary part
 * is read off the mirrored cosine, W[n/4-i].im = -W[i].re. */
void fxo_twiddle_r2(int n, fxo_cf32* W)
{
    int q = n / 4;
    W[0].re = 1.0f;
    W[0].im = 0.0f;
    for (int i = 1; i < q; i++) {
        W[i].re = (float)cos((double)i * 2.0 * M_PI / (double)n);
        W[q - i].im = -W[i].re;
    }
}

/* Twiddle W_n^e for e in [0, n/2): the second quarter is implied by the
 * symmetry stated in fft_1d_r2_modified.h:40-51,
 * W[i+n/4].re = W[i].im, W[i+n/4].im = -W[i].re. */
static fxo_cf32 tw_at(const fxo_cf32* W, int n, int e)
{
    int q = n / 4;
    fxo_cf32 w;
    if (e == 0) {
        /* n = 2 has an empty table; W^0 = 1 for every n */
        w.re = 1.0f;
        w.im = 0.0f;
    } else if (e < q) {
        w = W[e];
    } else {
        w.re = W[e - q].im;
        w.im = -W[e - q].re;
    }
    return w;
}

static int ilog2i(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

/* In-place forward complex FFT, X[k] = sum x[m] e^{-2 pi i m k/n}, natural
 * order output, unnormalized (fft_1d_r2_modified.h:16-36: radix-2 DIT
 * butterflies P = p + q*W, Q = p - q*W over log2(n) stages after a
 * bit-reversal of the input). */
void fxo_fft_r2_c2c(fxo_cf32* x, int n, const fxo_cf32* W)
{
    int lg = ilog2i(n);
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
        if (r > i) {
            fxo_cf32 t = x[i];
            x[i] = x[r];
            x[r] = t;
        }
    }
    for (int half = 1; half < n; half <<= 1) {
        int stride = n / (2 * half);
        for (int b = 0; b < n; b += 2 * half) {
            for (int k = 0; k < half; k++) {
                fxo_cf32 w = tw_at(W, n, k * stride);
                fxo_cf32 p = x[b + k], q = x[b + k + half];
                float qr = q.re * w.re - q.im * w.im;
                float qi = q.im * w.re + q.re * w.im;
                x[b + k].re = p.re + qr;
                x[b + k].im = p.im + qi;
                x[b + k + half].re = p.re - qr;
                x[b + k + half].im = p.im - qi;
            }
        }
    }
}

/* Real input through the complex FFT with zero imaginary part, the way the
 * SPU spectrometer feeds _fft_1d_r2 ({re,0} interleave,
 * spectrometer_spu.c:321-349), keeping bins 0..n/2-1 (Nyquist dropped). */
void fxo_rfft_r2(const float* x, int n, const fxo_cf32* W, fxo_cf32* X, fxo_cf32* work)
{
    for (int i = 0; i < n; i++) {
        work[i].re = x[i];
        work[i].im = 0.0f;
    }
    fxo_fft_r2_c2c(work, n, W);
    memcpy(X, work, (size_t)(n / 2) * sizeof(fxo_cf32));
}

void fxo_fft_f64(fxo_cf64* x, int n)
{
    int lg = ilog2i(n);
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
        if (r > i) {
            fxo_cf64 t = x[i];
            x[i] = x[r];
            x[r] = t;
        }
    }
    for (int half = 1; half < n; half <<= 1) {
        for (int k = 0; k < half; k++) {
            double ang = -M_PI * (double)k / (double)half;
            double wr = cos(ang), wi = sin(ang);
            for (int b = 0; b < n; b += 2 * half) {
                fxo_cf64 p = x[b + k], q = x[b + k + half];
                double qr = q.re * wr - q.im * wi;
                double qi = q.im * wr + q.re * wi;
                x[b + k].re = p.re + qr;
                x[b + k].im = p.im + qi;
                x[b + k + half].re = p.re - qr;
                x[b + k + half].im = p.im - qi;
            }
        }
    }
}

void fxo_dft_f64(const fxo_cf64* in, int n, fxo_cf64* out)
{
    for (int k = 0; k < n; k++) {
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < n; m++) {
            /* exact integer reduction of the angle index keeps the DFT exact */
            long e = ((long)m * (long)k) % n;
            double ang = -2.0 * M_PI * (double)e / (double)n;
            double c = cos(ang), s = sin(ang);
            sr += in[m].re * c - in[m].im * s;
            si += in[m].re * s + in[m].im * c;
        }
        out[k].re = sr;
        out[k].im = si;
    }
}

/* ================================================================== */
/* a4: fringe/delay rotation                                          */
/* ================================================================== */

/* phi(k,t) = 2 pi k tau(t)/n + fringe(t), tau(t) = d0 + d1 t,
 * fringe(t) = f0 + f1 t, evaluated and reduced to [-pi, pi] in FP64.
 * The DiFX task applies such a per-channel phasor with a complex multiply
 * (complexfracmult, cellspe-tasklib-dsp.cpp:671-673). */
double fxo_phase(const fxo_phase_model* m, int k, int n, long t)
{
    double tau = m->delay0 + m->delay_rate * (double)t;
    double cyc = (double)k * tau / (double)n + (m->fringe0 + m->fringe_rate * (double)t) / (2.0 * M_PI);
    cyc -= rint(cyc);
    return 2.0 * M_PI * cyc;
}

/* Y = X * e^{i phi} with the FP32 phasor (cosf, sinf of the reduced FP32
 * phase) and genericMul_32fc's arithmetic (architecture.h:270-277). */
void fxo_rotate(fxo_cf32* X, int nchan, int n, const fxo_phase_model* m, long t)
{
    for (int k = 0; k < nchan; k++) {
        float ph = (float)fxo_phase(m, k, n, t);
        fxo_cf32 w = { cosf(ph), sinf(ph) };
        fxo_cf32 s = X[k];
        X[k].re = (s.re * w.re) - (s.im * w.im);
        X[k].im = (s.im * w.re) + (s.re * w.im);
    }
}

void fxo_rotate_f64(fxo_cf64* X, int nchan, int n, const fxo_phase_model* m, long t)
{
    for (int k = 0; k < nchan; k++) {
        double ph = fxo_phase(m, k, n, t);
        double wr = cos(ph), wi = sin(ph);
        fxo_cf64 s = X[k];
        X[k].re = s.re * wr - s.im * wi;
        X[k].im = s.im * wr + s.re * wi;
    }
}

/* ================================================================== */
/* a6: complex MAC primitives and X-engine                             */
/* ================================================================== */

/* genericMul_32fc (architecture.h:270-277) */
void fxo_mul(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* out, int len)
{
    for (int i = 0; i < len; i++) {
        fxo_cf32 r;
        r.re = (a[i].re * b[i].re) - (a[i].im * b[i].im);
        r.im = (a[i].im * b[i].re) + (a[i].re * b[i].im);
        out[i] = r;
    }
}

/* genericAddProduct_32fc (architecture.h:405-424): acc += a*b, no conjugate.
 * This is what speAddProduct_32fc / CELLDSP_TASK_MAC computes. */
void fxo_addproduct(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* acc, int len)
{
    for (int i = 0; i < len; i++) {
        acc[i].re += (a[i].re * b[i].re) - (a[i].im * b[i].im);
        acc[i].im += (a[i].im * b[i].re) + (a[i].re * b[i].im);
    }
}

/* genericAddProduct_32fc_conj (architecture.h:427-434): acc += a*conj(b). */
void fxo_addproduct_conj(const fxo_cf32* a, const fxo_cf32* b, fxo_cf32* acc, int len)
{
    for (int i = 0; i < len; i++) {
        acc[i].re += (a[i].re * b[i].re) + (a[i].im * b[i].im);
        acc[i].im += (a[i].im * b[i].re) - (a[i].re * b[i].im);
    }
}

/* Canonical product index p(i,j) = i*I - i(i-1)/2 + (j-i), i <= j. */
size_t fxo_prod_index(int i, int j, int ninp)
{
    return (size_t)i * (size_t)ninp - (size_t)i * (size_t)(i - 1) / 2 + (size_t)(j - i);
}

size_t fxo_nprod(int ninp)
{
    return (size_t)ninp * (size_t)(ninp + 1) / 2;
}

void fxo_xengine_f32(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf32* vis)
{
    for (long t = 0; t < ntime; t++) {
        for (int c = 0; c < nchan; c++) {
            const fxo_cf32* row = X + ((size_t)c * (size_t)ntime + (size_t)t) * (size_t)ninp;
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++) {
                    fxo_cf32* acc = vis + fxo_prod_index(i, j, ninp) * (size_t)nchan + (size_t)c;
                    acc->re += (row[i].re * row[j].re) + (row[i].im * row[j].im);
                    acc->im += (row[i].im * row[j].re) - (row[i].re * row[j].im);
                }
        }
    }
}

/* fxo_xengine_f32 with the channels shared among nthreads workers: every
 * (product, channel) sum runs over t in the same order, so the results equal
 * the single-threaded ones bit for bit. */
void fxo_xengine_f32_mt(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf32* vis, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
    for (int c = 0; c < nchan; c++) {
        for (long t = 0; t < ntime; t++) {
            const fxo_cf32* row = X + ((size_t)c * (size_t)ntime + (size_t)t) * (size_t)ninp;
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++) {
                    fxo_cf32* acc = vis + fxo_prod_index(i, j, ninp) * (size_t)nchan + (size_t)c;
                    acc->re += (row[i].re * row[j].re) + (row[i].im * row[j].im);
                    acc->im += (row[i].im * row[j].re) - (row[i].re * row[j].im);
                }
        }
    }
}

void fxo_xengine_f64(const fxo_cf32* X, int nchan, long ntime, int ninp, fxo_cf64* vis)
{
    for (int c = 0; c < nchan; c++) {
        for (long t = 0; t < ntime; t++) {
            const fxo_cf32* row = X + ((size_t)c * (size_t)ntime + (size_t)t) * (size_t)ninp;
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++) {
                    fxo_cf64* acc = vis + fxo_prod_index(i, j, ninp) * (size_t)nchan + (size_t)c;
                    double ar = row[i].re, ai = row[i].im, br = row[j].re, bi = row[j].im;
                    acc->re += ar * br + ai * bi;
                    acc->im += ai * br - ar * bi;
                }
        }
    }
}

/* One spectrum of one input: unpack -> FFT -> rotate (FP32). */
static void f_one_f32(const uint32_t* words, int n, long t, const fxo_phase_model* pm,
                      const fxo_cf32* W, float* xs, fxo_cf32* X, fxo_cf32* work)
{
    size_t off = (size_t)t * (size_t)n;
    for (int s = 0; s < n; s++) {
        size_t g = off + (size_t)s;
        xs[s] = fxo_lut2bit[(words[g >> 4] >> (2 * (g & 15))) & 3u];
    }
    fxo_rfft_r2(xs, n, W, X, work);
    if (pm) fxo_rotate(X, n / 2, n, pm, t);
}

static void f_one_f64(const uint32_t* words, int n, long t, const fxo_phase_model* pm,
                      fxo_cf64* work)
{
    size_t off = (size_t)t * (size_t)n;
    for (int s = 0; s < n; s++) {
        size_t g = off + (size_t)s;
        work[s].re = fxo_lut2bit[(words[g >> 4] >> (2 * (g & 15))) & 3u];
        work[s].im = 0.0;
    }
    fxo_fft_f64(work, n);
    if (pm) fxo_rotate_f64(work, n / 2, n, pm, t);
}

void fxo_fstage_f32(const uint32_t* const* words, int ninp, int n, long t0, long nt,
                    const fxo_phase_model* pm, fxo_cf32* X)
{
    int C = n / 2;
    fxo_cf32* W = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)(n / 4 + 1));
    float* xs = (float*)malloc(sizeof(float) * (size_t)n);
    fxo_cf32* spec = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)C);
    fxo_cf32* work = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)n);
    fxo_twiddle_r2(n, W);
    for (long tt = 0; tt < nt; tt++)
        for (int i = 0; i < ninp; i++) {
            f_one_f32(words[i], n, t0 + tt, pm ? &pm[i] : NULL, W, xs, spec, work);
            for (int c = 0; c < C; c++)
                X[((size_t)c * (size_t)nt + (size_t)tt) * (size_t)ninp + (size_t)i] = spec[c];
        }
    free(W); free(xs); free(spec); free(work);
}

/* Time-sliced chain: each worker integrates a contiguous range of spectra
 * into its own accumulator (TaskDispatcher.cpp:116-137), and the
 * accumulators are summed in worker order (combineResults, :141-170). */
void fxo_corr_f32(const uint32_t* const* words, int ninp, int n, long ntime,
                  const fxo_phase_model* pm, fxo_cf32* vis, int nthreads)
{
    int C = n / 2;
    size_t P = fxo_nprod(ninp);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ntime) nthreads = (int)ntime;
    fxo_cf32* part = (fxo_cf32*)calloc((size_t)nthreads * P * (size_t)C, sizeof(fxo_cf32));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        long ta = ntime * tid / nthreads, tb = ntime * (tid + 1) / nthreads;
        fxo_cf32* acc = part + (size_t)tid * P * (size_t)C;
        fxo_cf32* W = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)(n / 4 + 1));
        float* xs = (float*)malloc(sizeof(float) * (size_t)n);
        fxo_cf32* spec = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)C * (size_t)ninp);
        fxo_cf32* work = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)n);
        fxo_twiddle_r2(n, W);
        for (long t = ta; t < tb; t++) {
            for (int i = 0; i < ninp; i++)
                f_one_f32(words[i], n, t, pm ? &pm[i] : NULL, W, xs, spec + (size_t)i * (size_t)C, work);
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++)
                    fxo_addproduct_conj(spec + (size_t)i * (size_t)C, spec + (size_t)j * (size_t)C,
                                        acc + fxo_prod_index(i, j, ninp) * (size_t)C, C);
        }
        free(W); free(xs); free(spec); free(work);
    }
    /* workers' sums added in worker order; the elements are independent */
    const long ne = (long)(P * (size_t)C);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long e = 0; e < ne; e++) {
        fxo_cf32 s = part[e];
        for (int w = 1; w < nthreads; w++) {
            s.re += part[(size_t)w * P * (size_t)C + (size_t)e].re;
            s.im += part[(size_t)w * P * (size_t)C + (size_t)e].im;
        }
        vis[e].re += s.re;
        vis[e].im += s.im;
    }
    free(part);
}

/* The X-engine of channelized data X[t][i][c] (c < nchan) the way the
 * generic core loops (TaskCoreIPP.cpp:456-487 per source pair,
 * genericAddProduct_32fc_conj over the channels, architecture.h:427-434),
 * time-sliced over nthreads workers with their own accumulators summed in
 * worker order (TaskDispatcher.cpp:116-170).  vis[p][c] += . */
void fxo_xcorr_tic_f32(const fxo_cf32* X, int ninp, int nchan, long ntime, fxo_cf32* vis, int nthreads)
{
    size_t P = fxo_nprod(ninp);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ntime) nthreads = (int)ntime;
    fxo_cf32* part = (fxo_cf32*)calloc((size_t)nthreads * P * (size_t)nchan, sizeof(fxo_cf32));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        long ta = ntime * tid / nthreads, tb = ntime * (tid + 1) / nthreads;
        fxo_cf32* acc = part + (size_t)tid * P * (size_t)nchan;
        for (long t = ta; t < tb; t++) {
            const fxo_cf32* row = X + (size_t)t * (size_t)ninp * (size_t)nchan;
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++)
                    fxo_addproduct_conj(row + (size_t)i * (size_t)nchan, row + (size_t)j * (size_t)nchan,
                                        acc + fxo_prod_index(i, j, ninp) * (size_t)nchan, nchan);
        }
    }
    const long ne = (long)(P * (size_t)nchan);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long e = 0; e < ne; e++) {
        fxo_cf32 s = part[e];
        for (int w = 1; w < nthreads; w++) {
            s.re += part[(size_t)w * P * (size_t)nchan + (size_t)e].re;
            s.im += part[(size_t)w * P * (size_t)nchan + (size_t)e].im;
        }
        vis[e].re += s.re;
        vis[e].im += s.im;
    }
    free(part);
}

void fxo_corr_f64(const uint32_t* const* words, int ninp, int n, long ntime,
                  const fxo_phase_model* pm, fxo_cf64* vis, int nthreads)
{
    int C = n / 2;
    size_t P = fxo_nprod(ninp);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ntime) nthreads = (int)ntime;
    fxo_cf64* part = (fxo_cf64*)calloc((size_t)nthreads * P * (size_t)C, sizeof(fxo_cf64));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        long ta = ntime * tid / nthreads, tb = ntime * (tid + 1) / nthreads;
        fxo_cf64* acc = part + (size_t)tid * P * (size_t)C;
        fxo_cf64* spec = (fxo_cf64*)malloc(sizeof(fxo_cf64) * (size_t)n * (size_t)ninp);
        for (long t = ta; t < tb; t++) {
            for (int i = 0; i < ninp; i++)
                f_one_f64(words[i], n, t, pm ? &pm[i] : NULL, spec + (size_t)i * (size_t)n);
            for (int i = 0; i < ninp; i++)
                for (int j = i; j < ninp; j++) {
                    const fxo_cf64* a = spec + (size_t)i * (size_t)n;
                    const fxo_cf64* b = spec + (size_t)j * (size_t)n;
                    fxo_cf64* v = acc + fxo_prod_index(i, j, ninp) * (size_t)C;
                    for (int c = 0; c < C; c++) {
                        v[c].re += a[c].re * b[c].re + a[c].im * b[c].im;
                        v[c].im += a[c].im * b[c].re - a[c].re * b[c].im;
                    }
                }
        }
        free(spec);
    }
    const long ne = (long)(P * (size_t)C);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long e = 0; e < ne; e++)
        for (int w = 0; w < nthreads; w++) {
            vis[e].re += part[(size_t)w * P * (size_t)C + (size_t)e].re;
            vis[e].im += part[(size_t)w * P * (size_t)C + (size_t)e].im;
        }
    free(part);
}

/* ================================================================== */
/* a8: spectrometer                                                   */
/* ================================================================== */

/* Window tables, evaluated in FP64 and rounded once to FP32.
 *  SIN2:    sin^2(pi i/(n-1))               (spectrometer_spu.c:297-312)
 *  COSINE:  cos(pi (i - n/2 + 1/2)/n)        (TaskCoreIPP.cpp:575-579)
 *  COSINE2: cos^2(pi (i - n/2 + 1/2)/n)      (TaskCoreIPP.cpp:581-586) */
void fxo_window(int kind, int n, float* w)
{
    for (int i = 0; i < n; i++) {
        double v = 1.0;
        if (kind == FXO_WIN_SIN2) {
            double s = sin(M_PI * (double)i / (double)(n - 1));
            v = s * s;
        } else if (kind == FXO_WIN_COSINE) {
            v = cos(M_PI * ((double)i - 0.5 * (double)n + 0.5) / (double)n);
        } else if (kind == FXO_WIN_COSINE2) {
            double c = cos(M_PI * ((double)i - 0.5 * (double)n + 0.5) / (double)n);
            v = c * c;
        } else if (kind == FXO_WIN_HAMMING) {
            v = 0.54 - 0.46 * cos(2.0 * M_PI * (double)i / (double)(n - 1));
        } else if (kind == FXO_WIN_HANN) {
            v = 0.5 - 0.5 * cos(2.0 * M_PI * (double)i / (double)(n - 1));
        }
        w[i] = (float)v;
    }
    if (kind == FXO_WIN_BLACKMAN) fxo_window_blackman(0.5, n, w);
}

void fxo_window_blackman(double alpha, int n, float* w)
{
    for (int i = 0; i < n; i++) {
        double x = 2.0 * M_PI * (double)i / (double)(n - 1);
        w[i] = (float)((alpha + 1.0) / 2.0 - 0.5 * cos(x) - 0.5 * alpha * cos(2.0 * x));
    }
}

/* Mean power spectrum (1/K) sum_k |X_k|^2, the 1/K applied once after the
 * accumulation as TaskCoreIPP does (TaskCoreIPP.cpp:124,437-447,489-493)
 * rather than the SPU kernel's pre-squaring scale (spectrometer_spu.c:269). */
void fxo_spectrometer_f32(const uint32_t* words, size_t ntime, int nchan, int n,
                          int integrate_ffts, int window_kind, float* out, int nthreads)
{
    fxo_spectrometer_ex_f32(words, ntime, nchan, n, 1, integrate_ffts, window_kind, 0, 0, out, nthreads);
}

/* Overlapped variant: FFT j of spectrum s starts at s*stride + j*hop,
 * hop = n/overlap; stride = K*hop for one unbroken run (spectrometer_spu.c:
 * 171-210) or (K-1)*hop + n when spectra do not share samples
 * (TaskCoreIPP.cpp:495-510). */
size_t fxo_spectrometer_ex_f32(const uint32_t* words, size_t ntime, int nchan, int n, int overlap,
                               int integrate_ffts, int window_kind, int continuous, int keep_nyquist,
                               float* out, int nthreads)
{
    int C = n / 2 + (keep_nyquist ? 1 : 0);
    size_t hop = (size_t)(n / overlap);
    size_t stride, nspec;
    if (continuous) {
        size_t nfft = ntime >= (size_t)n ? (ntime - (size_t)n) / hop + 1 : 0;
        nspec = nfft / (size_t)integrate_ffts;
        stride = (size_t)integrate_ffts * hop;
    } else {
        stride = (size_t)(integrate_ffts - 1) * hop + (size_t)n;
        nspec = ntime / stride;
    }
    if (!out) return nspec;
    float* win = (float*)malloc(sizeof(float) * (size_t)n);
    fxo_window(window_kind, n, win);
    int spw = 16 / nchan;
    long njobs = (long)nspec * nchan;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
#endif
    for (long job = 0; job < njobs; job++) {
        size_t s = (size_t)(job / nchan);
        int c = (int)(job % nchan);
        fxo_cf32* W = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)(n / 4 + 1));
        float* xs = (float*)malloc(sizeof(float) * (size_t)n);
        fxo_cf32* X = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)(n / 2));
        fxo_cf32* work = (fxo_cf32*)malloc(sizeof(fxo_cf32) * (size_t)n);
        float* acc = (float*)calloc((size_t)C, sizeof(float));
        fxo_twiddle_r2(n, W);
        for (int f = 0; f < integrate_ffts; f++) {
            size_t t0 = s * stride + (size_t)f * hop;
            for (int i = 0; i < n; i++) {
                size_t t = t0 + (size_t)i;
                uint32_t wd = words[t / (size_t)spw];
                int pos = (int)(t % (size_t)spw) * nchan + c;
                xs[i] = fxo_lut2bit[(wd >> (2 * pos)) & 3u] * win[i];
            }
            fxo_rfft_r2(xs, n, W, X, work);
            /* work holds the full n-point transform; bin n/2 is the Nyquist */
            for (int k = 0; k < C; k++)
                acc[k] += work[k].re * work[k].re + work[k].im * work[k].im;
        }
        float scale = 1.0f / (float)integrate_ffts;
        for (int k = 0; k < C; k++)
            out[(s * (size_t)nchan + (size_t)c) * (size_t)C + (size_t)k] = acc[k] * scale;
        free(W); free(xs); free(X); free(work); free(acc);
    }
    free(win);
    return nspec;
}

void fxo_spectrometer_f64(const uint32_t* words, size_t ntime, int nchan, int n,
                          int integrate_ffts, int window_kind, double* out)
{
    int C = n / 2;
    size_t nfft = ntime / (size_t)n;
    size_t nspec = nfft / (size_t)integrate_ffts;
    float* win = (float*)malloc(sizeof(float) * (size_t)n);
    fxo_cf64* work = (fxo_cf64*)malloc(sizeof(fxo_cf64) * (size_t)n);
    double* acc = (double*)malloc(sizeof(double) * (size_t)C);
    fxo_window(window_kind, n, win);
    int spw = 16 / nchan;
    for (size_t s = 0; s < nspec; s++)
        for (int c = 0; c < nchan; c++) {
            memset(acc, 0, sizeof(double) * (size_t)C);
            for (int f = 0; f < integrate_ffts; f++) {
                size_t t0 = (s * (size_t)integrate_ffts + (size_t)f) * (size_t)n;
                for (int i = 0; i < n; i++) {
                    size_t t = t0 + (size_t)i;
                    uint32_t wd = words[t / (size_t)spw];
                    int pos = (int)(t % (size_t)spw) * nchan + c;
                    work[i].re = (double)fxo_lut2bit[(wd >> (2 * pos)) & 3u] * (double)win[i];
                    work[i].im = 0.0;
                }
                fxo_fft_f64(work, n);
                for (int k = 0; k < C; k++)
                    acc[k] += work[k].re * work[k].re + work[k].im * work[k].im;
            }
            for (int k = 0; k < C; k++)
                out[(s * (size_t)nchan + (size_t)c) * (size_t)C + (size_t)k] = acc[k] / (double)integrate_ffts;
        }
    free(win); free(work); free(acc);
}

/* ================================================================== */
/* a5: phase-cal extraction (FP64 accumulators)                        */
/* ================================================================== */

/* PCal::gcd (PCal.cpp:117-129) */
long long fxo_gcd(long a, long b)
{
    for (;;) {
        a = a % b;
        if (a == 0) return b;
        b = b % a;
        if (b == 0) return a;
    }
}

/* PCal::getNew (PCal.cpp:92-112) and the derived constructors
 * (Trivial :183-191, Shifting :320-329, ImplicitShift :521-528). */
void fxo_pcal_config(double bandwidth_hz, double spacing_hz, int offset_hz, fxo_pcal_cfg* cfg)
{
    double fs = 2.0 * bandwidth_hz;
    memset(cfg, 0, sizeof(*cfg));
    cfg->fs_hz = fs;
    cfg->spacing_hz = spacing_hz;
    cfg->offset_hz = (double)offset_hz;
    if (offset_hz == 0) {
        cfg->kind = FXO_PCAL_TRIVIAL;
        cfg->nbins = (int)(fs / (double)fxo_gcd((long)fabs(spacing_hz), (long)fs));
        cfg->ntones = (int)floor(bandwidth_hz / spacing_hz);
    } else {
        int No = (int)(fs / (double)fxo_gcd(offset_hz, (long)fs));
        int Np = (int)(fs / (double)fxo_gcd((long)spacing_hz, (long)fs));
        cfg->ntones = (int)floor((bandwidth_hz - offset_hz) / spacing_hz) + 1;
        if (No % Np == 0) {
            cfg->kind = FXO_PCAL_IMPLICIT;
            cfg->nbins = (int)(fs / (double)fxo_gcd(abs(offset_hz), (long)fs));
        } else {
            cfg->kind = FXO_PCAL_SHIFTING;
            cfg->nbins = (int)(fs / (double)fxo_gcd((long)fabs(spacing_hz), (long)fs));
            cfg->rotatorlen = (int)(fs / (double)fxo_gcd(abs(offset_hz), (long)fs));
        }
    }
    /* tone bin selection (PCal.cpp:306-312, 505-512, 664-676) */
    cfg->step = (int)floor((double)cfg->nbins * (spacing_hz / fs));
    cfg->offset_bin = (cfg->kind == FXO_PCAL_IMPLICIT)
                          ? (int)floor((double)cfg->nbins * cfg->offset_hz / fs) : 0;
}

/* Fold len samples starting at absolute sample index `sampleoffset` into
 * acc[2*nbins].  Trivial/ImplicitShift fold real samples at bin
 * (n mod nbins) (PCal.cpp:265-288, 610-644); Shifting multiplies by
 * e^{-2 pi i f_off n/fs} (rotator period rotatorlen) first (PCal.cpp:424-479).
 * The reference accumulates in Ipp32f; here the accumulators are FP64. */
void fxo_pcal_extract(const fxo_pcal_cfg* cfg, const float* x, size_t len,
                      size_t sampleoffset, fxo_cf64* acc)
{
    int nb = cfg->nbins;
    if (cfg->kind == FXO_PCAL_SHIFTING) {
        double dphi = 2.0 * M_PI * (-cfg->offset_hz / cfg->fs_hz);
        for (size_t i = 0; i < len; i++) {
            size_t n = sampleoffset + i;
            double arg = dphi * (double)(n % (size_t)cfg->rotatorlen);
            acc[n % (size_t)nb].re += cos(arg) * (double)x[i];
            acc[n % (size_t)nb].im += sin(arg) * (double)x[i];
        }
    } else {
        for (size_t i = 0; i < len; i++) {
            size_t n = sampleoffset + i;
            acc[n % (size_t)nb].re += (double)x[i];
        }
    }
}

/* getFinalPCal: nbins-point forward DFT scaled by 1/nbins
 * (IPP_FFT_DIV_FWD_BY_N) evaluated at the tone bins offset_bin + t*step. */
void fxo_pcal_final(const fxo_pcal_cfg* cfg, const fxo_cf64* acc, fxo_cf64* tones)
{
    int nb = cfg->nbins;
    for (int t = 0; t < cfg->ntones; t++) {
        long k = (long)cfg->offset_bin + (long)t * cfg->step;
        if (k >= nb) {
            tones[t].re = tones[t].im = 0.0;
            continue;
        }
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < nb; b++) {
            long e = ((long)b * k) % nb;
            double ang = -2.0 * M_PI * (double)e / (double)nb;
            double c = cos(ang), s = sin(ang);
            sr += acc[b].re * c - acc[b].im * s;
            si += acc[b].re * s + acc[b].im * c;
        }
        tones[t].re = sr / (double)nb;
        tones[t].im = si / (double)nb;
    }
}

/* PCal::extractAndIntegrate_reference (PCal.cpp:139-175): the slow direct
 * method, FP64: rotate every sample by e^{-2 pi i f_off n/fs}, fold into
 * 2*ntones bins, DFT with 1/N, keep the first ntones bins. */
void fxo_pcal_reference(const fxo_pcal_cfg* cfg, const float* x, size_t len,
                        size_t sampleoffset, fxo_cf64* tones)
{
    int nb = 2 * cfg->ntones;
    long period = (long)(cfg->fs_hz / (double)fxo_gcd((long)cfg->offset_hz, (long)cfg->fs_hz));
    double dphi = 2.0 * M_PI * (-cfg->offset_hz / cfg->fs_hz);
    fxo_cf64* acc = (fxo_cf64*)calloc((size_t)nb, sizeof(fxo_cf64));
    for (size_t i = 0; i < len; i++) {
        size_t no = i + sampleoffset;
        double phi = dphi * (double)(no % (size_t)period);
        acc[no % (size_t)nb].re += cos(phi) * (double)x[i];
        acc[no % (size_t)nb].im += sin(phi) * (double)x[i];
    }
    for (int t = 0; t < cfg->ntones; t++) {
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < nb; b++) {
            long e = ((long)b * t) % nb;
            double ang = -2.0 * M_PI * (double)e / (double)nb;
            double c = cos(ang), s = sin(ang);
            sr += acc[b].re * c - acc[b].im * s;
            si += acc[b].re * s + acc[b].im * c;
        }
        tones[t].re = sr / (double)nb;
        tones[t].im = si / (double)nb;
    }
    free(acc);
}

/* TaskCoreIPP::extract_PCal (TaskCoreIPP.cpp:633-693) with the rotator of
 * TaskCoreIPP.cpp:129-136, accumulated in FP64.  points must be a multiple of
 * rotatorlen (swspectrometer.cpp:185-190 disables the feature otherwise). */
void fxo_tcpcal_fold(const float* x, int points, int tonebins, int rotatorlen, double dphi,
                     fxo_cf64* acc)
{
    int pulses = rotatorlen / tonebins;
    for (int n = 0; n + rotatorlen <= points; n += rotatorlen) {
        for (int p = 0; p < pulses; p++) {
            for (int b = 0; b < tonebins; b++) {
                int r = p * tonebins + b;
                double arg = dphi * (double)r;
                double v = (double)x[n + r];
                acc[b].re += v * cos(arg);
                acc[b].im += v * sin(arg);
            }
        }
    }
}

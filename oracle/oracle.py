"""ctypes binding of the CPU oracle (oracle/liboracle_fx.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP F-X engine.  Imported by
tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never
by product code.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FX_ORACLE_LIB names another build of the restatement (the sanitizer build)
LIB_PATH = os.environ.get("FX_ORACLE_LIB") or os.path.join(HERE, "liboracle_fx.so")

LUT = np.array([-3.3359, -1.0, 1.0, 3.3359], dtype=np.float32)
WIN_NONE, WIN_SIN2, WIN_COSINE, WIN_COSINE2, WIN_HAMMING, WIN_BLACKMAN, WIN_HANN = range(7)
PCAL_TRIVIAL, PCAL_IMPLICIT, PCAL_SHIFTING = 0, 1, 2


class PhaseModel(C.Structure):
    _fields_ = [("delay0", C.c_double), ("delay_rate", C.c_double),
                ("fringe0", C.c_double), ("fringe_rate", C.c_double)]


class PCalCfg(C.Structure):
    _fields_ = [("kind", C.c_int), ("fs_hz", C.c_double), ("spacing_hz", C.c_double),
                ("offset_hz", C.c_double), ("nbins", C.c_int), ("ntones", C.c_int),
                ("rotatorlen", C.c_int), ("step", C.c_int), ("offset_bin", C.c_int)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle_fx.so"])


def build_native(out_path):
    """Build the restatement -O3 -march=native for this host at out_path (the
    benchmark's CPU baseline) and make it the library every call below uses."""
    global _lib
    subprocess.check_call(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out_path}"])
    _lib = None
    return load(out_path)


def load(path=None):
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = LIB_PATH
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    vp, sz, i, l, d, u64 = C.c_void_p, C.c_size_t, C.c_int, C.c_long, C.c_double, C.c_uint64
    sigs = {
        "fxo_mix64": (u64, [u64]),
        "fxo_uniform": (d, [u64, u64]),
        "fxo_gauss": (d, [u64, u64]),
        "fxo_stream_key": (u64, [u64, u64]),
        "fxo_quantize2": (i, [d, d]),
        "fxo_pack2b": (None, [vp, sz, vp]),
        "fxo_gen_station": (None, [u64, i, d, l, d, d, d, sz, vp]),
        "fxo_gen_multichannel": (None, [u64, i, vp, vp, d, sz, vp]),
        "fxo_gen_channelized_codes": (None, [u64, i, l, i, d, d, vp]),
        "fxo_unpack2b_1ch": (None, [vp, sz, vp]),
        "fxo_unpack2b_nch": (None, [vp, sz, i, vp]),
        "fxo_twiddle_r2": (None, [i, vp]),
        "fxo_fft_r2_c2c": (None, [vp, i, vp]),
        "fxo_rfft_r2": (None, [vp, i, vp, vp, vp]),
        "fxo_fft_f64": (None, [vp, i]),
        "fxo_dft_f64": (None, [vp, i, vp]),
        "fxo_phase": (d, [vp, i, i, l]),
        "fxo_rotate": (None, [vp, i, i, vp, l]),
        "fxo_mul": (None, [vp, vp, vp, i]),
        "fxo_addproduct": (None, [vp, vp, vp, i]),
        "fxo_addproduct_conj": (None, [vp, vp, vp, i]),
        "fxo_prod_index": (sz, [i, i, i]),
        "fxo_nprod": (sz, [i]),
        "fxo_xengine_f32": (None, [vp, i, l, i, vp]),
        "fxo_xengine_f32_mt": (None, [vp, i, l, i, vp, i]),
        "fxo_xcorr_tic_f32": (None, [vp, i, i, l, vp, i]),
        "fxo_xengine_f64": (None, [vp, i, l, i, vp]),
        "fxo_corr_f32": (None, [vp, i, i, l, vp, vp, i]),
        "fxo_corr_f64": (None, [vp, i, i, l, vp, vp, i]),
        "fxo_fstage_f32": (None, [vp, i, i, l, l, vp, vp]),
        "fxo_window": (None, [i, i, vp]),
        "fxo_spectrometer_f32": (None, [vp, sz, i, i, i, i, vp, i]),
        "fxo_spectrometer_f64": (None, [vp, sz, i, i, i, i, vp]),
        "fxo_spectrometer_ex_f32": (sz, [vp, sz, i, i, i, i, i, i, i, vp, i]),
        "fxo_window_blackman": (None, [d, i, vp]),
        "fxo_tcpcal_fold": (None, [vp, i, i, i, d, vp]),
        "fxo_gcd": (C.c_longlong, [l, l]),
        "fxo_pcal_config": (None, [d, d, i, C.POINTER(PCalCfg)]),
        "fxo_pcal_extract": (None, [C.POINTER(PCalCfg), vp, sz, sz, vp]),
        "fxo_pcal_final": (None, [C.POINTER(PCalCfg), vp, vp]),
        "fxo_pcal_reference": (None, [C.POINTER(PCalCfg), vp, sz, sz, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- generators -----------------------------------------------------------
def gen_station(seed, station, rho, delay, tone_amp, tone_freq, thresh, nsamp):
    w = np.zeros((nsamp + 15) // 16, dtype=np.uint32)
    load().fxo_gen_station(seed, station, rho, delay, tone_amp, tone_freq, thresh, nsamp, _p(w))
    return w


def gen_multichannel(seed, nchan, amp, freq, thresh, ntime):
    spw = 16 // nchan
    w = np.zeros((ntime + spw - 1) // spw, dtype=np.uint32)
    a = np.ascontiguousarray(amp, dtype=np.float64)
    f = np.ascontiguousarray(freq, dtype=np.float64)
    load().fxo_gen_multichannel(seed, nchan, _p(a), _p(f), thresh, ntime, _p(w))
    return w


def gen_channelized_codes(seed, nchan, ntime, ninp, rho, thresh):
    c = np.zeros(nchan * ntime * ninp, dtype=np.uint8)
    load().fxo_gen_channelized_codes(seed, nchan, ntime, ninp, rho, thresh, _p(c))
    return c.reshape(nchan, ntime, ninp)


def pack2b(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    w = np.zeros((codes.size + 15) // 16, dtype=np.uint32)
    load().fxo_pack2b(_p(codes), codes.size, _p(w))
    return w


# ---- a1 -----------------------------------------------------------------
def unpack_1ch(words, nsamp):
    out = np.zeros(nsamp, dtype=np.float32)
    load().fxo_unpack2b_1ch(_p(np.ascontiguousarray(words, np.uint32)), nsamp, _p(out))
    return out


def unpack_nch(words, ntime, nchan):
    out = np.zeros(nchan * ntime, dtype=np.float32)
    load().fxo_unpack2b_nch(_p(np.ascontiguousarray(words, np.uint32)), ntime, nchan, _p(out))
    return out.reshape(nchan, ntime)


# ---- a2/a3 --------------------------------------------------------------
def twiddles(n):
    W = np.zeros(n // 4 + 1, dtype=np.complex64)
    load().fxo_twiddle_r2(n, _p(W))
    return W[: n // 4]


def fft_r2(x):
    """Forward complex FFT of x (complex64), radix-2 FP32 restatement."""
    x = np.array(x, dtype=np.complex64, copy=True)
    n = x.size
    W = np.zeros(max(n // 4, 1) + 1, dtype=np.complex64)
    load().fxo_twiddle_r2(n, _p(W))
    load().fxo_fft_r2_c2c(_p(x), n, _p(W))
    return x


def rfft_r2(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.size
    W = np.zeros(n // 4 + 1, dtype=np.complex64)
    load().fxo_twiddle_r2(n, _p(W))
    X = np.zeros(n // 2, dtype=np.complex64)
    work = np.zeros(n, dtype=np.complex64)
    load().fxo_rfft_r2(_p(x), n, _p(W), _p(X), _p(work))
    return X


def fft_f64(x):
    x = np.array(x, dtype=np.complex128, copy=True)
    load().fxo_fft_f64(_p(x), x.size)
    return x


def dft_f64(x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    load().fxo_dft_f64(_p(x), x.size, _p(out))
    return out


# ---- a4 / a6 --------------------------------------------------------------
def phase_models(models):
    arr = (PhaseModel * len(models))(*[PhaseModel(*m) for m in models])
    return arr


def phase(model, k, n, t):
    pm = PhaseModel(*model)
    return load().fxo_phase(C.byref(pm), k, n, t)


def rotate(X, n, model, t):
    X = np.array(X, dtype=np.complex64, copy=True)
    pm = PhaseModel(*model)
    load().fxo_rotate(_p(X), X.size, n, C.byref(pm), t)
    return X


def addproduct(a, b, acc, conj):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    acc = np.array(acc, dtype=np.complex64, copy=True)
    fn = load().fxo_addproduct_conj if conj else load().fxo_addproduct
    fn(_p(a), _p(b), _p(acc), a.size)
    return acc


def mul(a, b):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    out = np.zeros_like(a)
    load().fxo_mul(_p(a), _p(b), _p(out), a.size)
    return out


def prod_index(i, j, ninp):
    return load().fxo_prod_index(i, j, ninp)


def xengine_f32(X):
    """X[c][t][i] complex64 -> vis[p][c] complex64."""
    X = np.ascontiguousarray(X, np.complex64)
    C_, T, I = X.shape
    vis = np.zeros((I * (I + 1) // 2, C_), dtype=np.complex64)
    load().fxo_xengine_f32(_p(X), C_, T, I, _p(vis))
    return vis


def xengine_f32_mt(X, nthreads):
    """xengine_f32 with the channels shared among nthreads threads (same sums)."""
    X = np.ascontiguousarray(X, np.complex64)
    C_, T, I = X.shape
    vis = np.zeros((I * (I + 1) // 2, C_), dtype=np.complex64)
    load().fxo_xengine_f32_mt(_p(X), C_, T, I, _p(vis), nthreads)
    return vis


def xcorr_tic_f32(X, nthreads):
    """X[t][i][c] complex64 -> vis[p][c]: the reference's loop order
    (genericAddProduct_32fc_conj over the channels per input pair),
    time-sliced over nthreads workers summed in worker order."""
    X = np.ascontiguousarray(X, np.complex64)
    T, I, C_ = X.shape
    vis = np.zeros((I * (I + 1) // 2, C_), dtype=np.complex64)
    load().fxo_xcorr_tic_f32(_p(X), I, C_, T, _p(vis), nthreads)
    return vis


def xengine_f64(X):
    X = np.ascontiguousarray(X, np.complex64)
    C_, T, I = X.shape
    vis = np.zeros((I * (I + 1) // 2, C_), dtype=np.complex128)
    load().fxo_xengine_f64(_p(X), C_, T, I, _p(vis))
    return vis


def round_int_f32(n, e):
    """n * 2^e for an integer n of any size, rounded to the nearest FP32 with
    ties to even (normal range; test-only)."""
    if n == 0:
        return 0.0
    a, sgn = abs(n), -1 if n < 0 else 1
    sh = a.bit_length() - 24
    if sh > 0:
        q, rem = a >> sh, a & ((1 << sh) - 1)
        half = 1 << (sh - 1)
        if rem > half or (rem == half and (q & 1)):
            q += 1
        a, e = q, e + sh
    return sgn * float(a) * 2.0 ** e


def xengine_2bit_exact_f32(codes):
    """Visibilities of 2-bit channelized codes (codes[c][t][i], byte re | im << 2,
    levels LUT) as the exact sums rounded once to FP32.  Returns [p][c]
    complex64 (p over i <= j): the exact counterpart of xengine_f64 on the
    level values (test-only).

    Each level times 2^22 is an integer, 2^22 s + d m with s its sign, m = s
    on the outer levels (|x| = H = h 2^-22, FP32 H has 22 fraction bits) and
    0 on the inner ones, and d = h - 2^22.  A product sum is then
    2^-44 (2^44 SS + 2^22 d (SM + MS) + d^2 MM) with SS, SM, MS, MM integer
    sums of +-1 and 0 (FP64 matrix products, exact below 2^53), combined in
    int64 (|value| < 2^62 for up to 8192 spectra) and rounded once by the
    int64 -> FP32 conversion (round to nearest, ties to even)."""
    codes = np.asarray(codes, np.uint8)
    C, T, I = codes.shape
    assert T <= 8192, "int64 combination holds up to 8192 spectra"
    h = int(np.ldexp(np.float64(LUT[3]), 22))
    assert float(h) == np.ldexp(np.float64(LUT[3]), 22) and LUT[2] == 1.0
    d = np.int64(h - (1 << 22))
    iu = np.triu_indices(I)
    out = np.empty((len(iu[0]), C), np.complex64)
    scale = np.float32(2.0 ** -44)

    def parts(c):
        s = np.where(c >= 2, 1.0, -1.0)
        return s, np.where((c == 0) | (c == 3), s, 0.0)

    def combine(ss, sm, ms, mm):
        # all four are integer-valued FP64 arrays
        n = (ss.astype(np.int64) << 44) + (sm + ms).astype(np.int64) * d * (1 << 22) + mm.astype(np.int64) * (d * d)
        return n.astype(np.float32) * scale

    for c in range(C):
        sr, mr = parts((codes[c] & 3).T)           # [i][t]
        si, mi = parts(((codes[c] >> 2) & 3).T)
        A = np.concatenate([sr, si], 1)           # re part first, then im along t
        B = np.concatenate([mr, mi], 1)
        Ax = np.concatenate([si, -sr], 1)         # (im, -re): x_i conj(x_j) imaginary part
        Bx = np.concatenate([mi, -mr], 1)
        re = combine(A @ A.T, A @ B.T, B @ A.T, B @ B.T)
        im = combine(Ax @ A.T, Ax @ B.T, Bx @ A.T, Bx @ B.T)
        out[:, c] = re[iu] + 1j * im[iu]
    return out


def _word_ptrs(streams):
    arrs = [np.ascontiguousarray(s, np.uint32) for s in streams]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def corr_f32(streams, n, ntime, models=None, nthreads=1):
    arrs, ptrs = _word_ptrs(streams)
    I = len(arrs)
    vis = np.zeros((I * (I + 1) // 2, n // 2), dtype=np.complex64)
    pm = phase_models(models) if models is not None else None
    load().fxo_corr_f32(C.cast(ptrs, C.c_void_p), I, n, ntime,
                        None if pm is None else C.cast(pm, C.c_void_p), _p(vis), nthreads)
    return vis


def corr_f64(streams, n, ntime, models=None, nthreads=1):
    arrs, ptrs = _word_ptrs(streams)
    I = len(arrs)
    vis = np.zeros((I * (I + 1) // 2, n // 2), dtype=np.complex128)
    pm = phase_models(models) if models is not None else None
    load().fxo_corr_f64(C.cast(ptrs, C.c_void_p), I, n, ntime,
                        None if pm is None else C.cast(pm, C.c_void_p), _p(vis), nthreads)
    return vis


def fstage_f32(streams, n, t0, nt, models=None):
    arrs, ptrs = _word_ptrs(streams)
    I = len(arrs)
    X = np.zeros((n // 2, nt, I), dtype=np.complex64)
    pm = phase_models(models) if models is not None else None
    load().fxo_fstage_f32(C.cast(ptrs, C.c_void_p), I, n, t0, nt,
                          None if pm is None else C.cast(pm, C.c_void_p), _p(X))
    return X


# ---- a8 -----------------------------------------------------------------
def window(kind, n):
    w = np.zeros(n, dtype=np.float32)
    load().fxo_window(kind, n, _p(w))
    return w


def spectrometer_f32(words, ntime, nchan, n, integrate_ffts, window_kind, nthreads=1):
    nspec = (ntime // n) // integrate_ffts
    out = np.zeros((nspec, nchan, n // 2), dtype=np.float32)
    load().fxo_spectrometer_f32(_p(np.ascontiguousarray(words, np.uint32)), ntime, nchan, n,
                                integrate_ffts, window_kind, _p(out), nthreads)
    return out


def spectrometer_ex_f32(words, ntime, nchan, n, overlap, integrate_ffts, window_kind, continuous,
                        keep_nyquist, nthreads=1):
    """Overlapped spectrometer; returns out[s][c][n/2 (+1)]."""
    w = np.ascontiguousarray(words, np.uint32)
    lib = load()
    nspec = lib.fxo_spectrometer_ex_f32(_p(w), ntime, nchan, n, overlap, integrate_ffts, window_kind,
                                        int(continuous), int(keep_nyquist), None, nthreads)
    out = np.zeros((nspec, nchan, n // 2 + int(bool(keep_nyquist))), dtype=np.float32)
    if nspec:
        lib.fxo_spectrometer_ex_f32(_p(w), ntime, nchan, n, overlap, integrate_ffts, window_kind,
                                    int(continuous), int(keep_nyquist), _p(out), nthreads)
    return out


def window_blackman(alpha, n):
    w = np.zeros(n, dtype=np.float32)
    load().fxo_window_blackman(float(alpha), n, _p(w))
    return w


def tcpcal_fold(x, tonebins, rotatorlen, dphi, acc=None):
    """IPP-core phase-cal of one FFT input x (len multiple of rotatorlen)."""
    x = np.ascontiguousarray(x, np.float32)
    if acc is None:
        acc = np.zeros(tonebins, dtype=np.complex128)
    load().fxo_tcpcal_fold(_p(x), x.size, tonebins, rotatorlen, float(dphi), _p(acc))
    return acc


def spectrometer_f64(words, ntime, nchan, n, integrate_ffts, window_kind):
    nspec = (ntime // n) // integrate_ffts
    out = np.zeros((nspec, nchan, n // 2), dtype=np.float64)
    load().fxo_spectrometer_f64(_p(np.ascontiguousarray(words, np.uint32)), ntime, nchan, n,
                                integrate_ffts, window_kind, _p(out))
    return out


# ---- a5 -----------------------------------------------------------------
def pcal_config(bandwidth_hz, spacing_hz, offset_hz):
    cfg = PCalCfg()
    load().fxo_pcal_config(bandwidth_hz, spacing_hz, int(offset_hz), C.byref(cfg))
    return cfg


def pcal_extract(cfg, x, sampleoffset=0, acc=None):
    x = np.ascontiguousarray(x, np.float32)
    if acc is None:
        acc = np.zeros(2 * cfg.nbins, dtype=np.complex128)
    load().fxo_pcal_extract(C.byref(cfg), _p(x), x.size, sampleoffset, _p(acc))
    return acc


def pcal_final(cfg, acc):
    out = np.zeros(max(cfg.ntones, 1), dtype=np.complex128)
    load().fxo_pcal_final(C.byref(cfg), _p(acc), _p(out))
    return out[: cfg.ntones]


def pcal_reference(cfg, x, sampleoffset=0):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(max(cfg.ntones, 1), dtype=np.complex128)
    load().fxo_pcal_reference(C.byref(cfg), _p(x), x.size, sampleoffset, _p(out))
    return out[: cfg.ntones]


# ---- visibility products (zerocorr.c, m5pcal.c) ---------------------------
def lags(vis, scale):
    """Lag spectra of visibility rows (zerocorr.c:391,633-645): FFTW_BACKWARD
    over the channels (unnormalized, sign +) times scale, natural order, as
    conj(FFT(conj v)) with the radix-2 FP32 restatement."""
    vis = np.atleast_2d(np.asarray(vis, np.complex64))
    out = np.empty_like(vis)
    for r in range(vis.shape[0]):
        out[r] = np.conj(fft_r2(np.conj(vis[r]))) * np.float32(scale)
    return out


def lags_f64(vis, scale):
    """The same in FP64 (exact-twiddle transform), for pinning."""
    vis = np.atleast_2d(np.asarray(vis, np.complex128))
    return np.stack([np.conj(fft_f64(np.conj(v))) * scale for v in vis])


def amp_phase(z, scale):
    """zerocorr.c:628-630,639-643: x, y = scale * (re, im); sqrt(x^2 + y^2),
    atan2(y, x), in FP64."""
    z = np.asarray(z, np.complex128) * float(scale)
    return np.sqrt(z.real ** 2 + z.imag ** 2), np.arctan2(z.imag, z.real)


def pcal_delay(tones, tone0_hz, tone_step_hz, band_center_hz, band_valid_hz):
    """m5pcal calcDelay (m5pcal.c:94-150) on one input's complex tones at
    tone0 + t*step Hz, phases atan2(im, re).  The reference slope comes from
    the tones (n/2, n/2+1); with two tones the reference reads past the end
    of its arrays, here it uses tones (0, 1).  Returns ns."""
    z = np.asarray(tones, np.complex128)
    nt = z.size
    if nt < 2:
        return 0.0
    ph = [float(np.arctan2(v.imag, v.real)) for v in z]
    f = [(tone0_hz + t * tone_step_hz) * 1e-6 for t in range(nt)]
    bc, bv = band_center_hz * 1e-6, band_valid_hz * 1e-6
    a = nt // 2 if nt >= 3 else 0
    ref = ph[a + 1] - ph[a]
    s, n = 0.0, 0
    for i in range(1, nt):
        if abs(f[i - 1] - bc) > bv or abs(f[i] - bc) > bv:
            continue
        d = ph[i] - ph[i - 1]
        if d - ref > np.pi:
            d -= 2.0 * np.pi
        if ref - d > np.pi:
            d += 2.0 * np.pi
        s += d / (f[i] - f[i - 1])
        n += 1
    if n < 1:
        return 0.0
    window = abs(500.0 / (tone_step_hz * 1e-6))
    delay = 1000.0 * (s / n) / (2.0 * np.pi)
    while delay > window:
        delay -= 2.0 * window
    while delay < -window:
        delay += 2.0 * window
    return delay


def spectrometer_xpol_f64(words_a, nchan_a, chan_a, words_b, nchan_b, chan_b, ntime, n, overlap,
                          integrate_ffts, window_kind, continuous=False, valid_a=None, valid_b=None):
    """Cross spectrum of two 2-bit channels (TaskCoreIPP.cpp:456-487) in FP64:
    the same windowed, overlapped frames as the spectrometer; per spectrum
    (1/K) sum X_a conj(X_b) over bins 0..n/2, with DC and Nyquist the product
    of the real parts (imaginary part 0).  valid_a/valid_b: optional per-word
    flags; samples of a word flagged 0 are zero (mark5access decodes blanked
    data as 0).  Returns [nspec][n/2+1] complex128."""
    hop = n // overlap
    K = integrate_ffts
    if continuous:
        nfft = (ntime - n) // hop + 1 if ntime >= n else 0
        nspec, stride = nfft // K, K * hop
    else:
        stride = (K - 1) * hop + n
        nspec = ntime // stride
    xa = unpack_nch(words_a, ntime, nchan_a)[chan_a].astype(np.float64)
    xb = unpack_nch(words_b, ntime, nchan_b)[chan_b].astype(np.float64)
    for x, v, nch in ((xa, valid_a, nchan_a), (xb, valid_b, nchan_b)):
        if v is not None:
            word_of_t = (np.arange(ntime) * nch) // 16     # word of each time step (same for every channel)
            x[~np.asarray(v, bool)[word_of_t]] = 0.0
    w = window(window_kind, n).astype(np.float64)
    out = np.zeros((nspec, n // 2 + 1), dtype=np.complex128)
    for s in range(nspec):
        t0 = s * stride + np.arange(K)[:, None] * hop + np.arange(n)[None, :]
        A = np.fft.rfft(xa[t0] * w, axis=1)
        B = np.fft.rfft(xb[t0] * w, axis=1)
        acc = (A * np.conj(B)).sum(axis=0)
        acc[0] = (A[:, 0].real * B[:, 0].real).sum()
        acc[-1] = (A[:, -1].real * B[:, -1].real).sum()
        out[s] = acc / K
    return out
